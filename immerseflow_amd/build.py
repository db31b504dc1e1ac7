"""Build libimmerseflow_amd.so (HIP, gfx950) and the immerseflow CLI in-tree.

    python -m immerseflow_amd.build        # or: python immerseflow_amd/build.py

Every HIP source is compiled for gfx950 without FP contraction (the FP64
path evaluates the reference kernels' expressions exactly as written) and
with packed-FP32 instructions disabled.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libimmerseflow_amd.so")
CLI = os.path.join(PKG, "immerseflow")

HIP_SOURCES = ["kernels.hip", "sweep_stream.hip", "sweep_reg.hip", "geometry.hip", "engine.hip", "compat.hip"]
CPP_SOURCES = ["io.cpp", "abi.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++20", "-ffp-contract=off",
          "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
          "-Wall", "-Wno-unused-function", "-Wno-unused-parameter"]
HIP_FLAGS = ["-x", "hip", "--offload-arch=gfx950",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def rocm_include():
    """ROCm's include directory (HIP runtime headers for the host-only C++)."""
    return os.path.join(os.path.dirname(os.path.dirname(os.path.realpath(hipcc()))), "include")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n%s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    jobs = []
    for s in HIP_SOURCES:
        jobs.append([cc] + COMMON + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o",
                                               os.path.join(BUILD, s + ".o")])
    for s in CPP_SOURCES:  # host-only C++ (HIP runtime API, no device code)
        jobs.append([cc] + COMMON + ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I" + rocm_include(), "-c", os.path.join(CSRC, s), "-o",
                                     os.path.join(BUILD, s + ".o")])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [os.path.join(BUILD, s + ".o") for s in HIP_SOURCES + CPP_SOURCES]
    tmp = LIB + ".tmp"
    _run([cc, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    _run([cc] + COMMON + HIP_FLAGS + [os.path.join(CSRC, "hbm_copy.hip"), "-o",
                                     os.path.join(PKG, "hbm_copy")])
    cli_src = os.path.join(CSRC, "cli.cpp")
    if os.path.exists(cli_src):
        _run([cc] + COMMON + ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I" + rocm_include(), cli_src, "-x", "none", "-o", CLI,
                              "-L" + PKG, "-limmerseflow_amd", "-Wl,-rpath,$ORIGIN"])
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
