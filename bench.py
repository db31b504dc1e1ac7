#!/usr/bin/env python3
"""Throughput of the fractional step on one MI355X.

Headline workload (the largest single-GPU configuration): eight seeded star
bodies in a 20 x 20 box, 4096 x 4096 uniform grid, Re = 1000, dt = 0.5 h,
Poisson stop max|div| <= 1e-7, mixed precision (FP32 Poisson error sweeps,
FP64 residual and defect correction, FP64 velocity passes, faces and
divergence).  A step is one full fractional step (predictor, divergence
RHS, Poisson, projection, divergence check) on device-resident fields.
Step 1 runs untimed and is followed by the autotuner; then --warmup - 1
more untimed steps and --steps timed steps.

value = interior cells x steps / wall seconds (Mcell-steps/s).

roofline: the dominant kernel is the Poisson sweep.  Temporal blocking
keeps several red-black sweeps per launch on chip, so the sweep is bound by
vector-ALU issue, not by HBM: achieved = FP32 operations of the cell
updates the launch executes (11 per update: one multiply and five fused
multiply-adds; the temporal-blocking halo recompute included) / mean
launch time, from device time stamps the library records on its stream
around each chunk of sweeps; peak = the FP32 rate without packed
instructions (the published rate counts packed FP32).  The measured HBM
traffic and VALU instruction count of the same launch (rocprofv3
FETCH_SIZE / WRITE_SIZE / SQ_INSTS_VALU, profiles/) give the HBM fraction
and the measured issue fraction beside it.

cpu_baseline: the C oracle (the same per-cell routines in FP64, OpenMP) on
the host's cores, on a bounded sample of the headline workload: one step
from the GPU's state with the Poisson loop capped, extrapolated to the
GPU's sweep count.  The 256^2 cylinder (the CPU-runnable workload) runs on
the CPU in full (500 steps) beside it.

Beside the headline: the 2048 x 1024 ellipse (all 2000 steps, mixed; FP64
and the FP32-velocity variant with their errors), the 256^2 cylinder on the
GPU (all 500 steps, FP64), the 256-case batch (all 200 steps, against the
plain FP64 single-case path on all 256 cases over a window of steps), a
prefix of the serving trace at 50 req/s and at a sub-saturating rate, and
the line-SOR Poisson option.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# MI355X published FP32 vector peak (TFLOP/s); it counts packed FP32 (two
# lanes per operation).  Without packed instructions the issue limit is
# half of it: 256 CUs x 64 lanes x 2 FLOP (FMA) x 2.4 GHz.
FP32_PEAK_TFLOPS = 157.3
FP32_NONPACKED_TFLOPS = 256 * 64 * 2 * 2.4e9 / 1e12
# FP32 red-black update, FLOP and vector instructions (the lane shift folded
# in): the general form (1 multiply + 5 FMA) and the form of a grid with one
# coefficient value per axis (1 add + 4 FMA)
FLOP_PER_UPDATE = {False: 11, True: 9}
VALU_PER_UPDATE = {False: 6, True: 5}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--full", action="store_true",
                    help="also run the headline workload for all 1000 steps and the FP64 single-case path "
                         "on all 256 batch cases for all 200 steps (long)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch the Poisson sweep kernels one by one instead of replaying a hipGraph")
    ap.add_argument("--no-persistent", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-line", action="store_true")
    ap.add_argument("--line-cases", default="c1,c2", help="line-SOR workloads: c1 (256^2), c2 (2048x1024)")
    ap.add_argument("--c3-tune", default="",
                    help="kernel,tile_y,sweeps_per_launch,check_every for the headline instead of the autotuner "
                         "(profiling runs: the autotuner's choice between close shapes can change under a profiler)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--c2-steps", type=int, default=2000)
    ap.add_argument("--c2-fp64-steps", type=int, default=3)
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--batch-window", type=int, default=2)
    ap.add_argument("--serve-requests", type=int, default=16)
    ap.add_argument("--cpu-sweeps", type=int, default=100,
                    help="Poisson sweeps of each capped oracle step in the CPU sample")
    return ap.parse_args()


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def kernel_name(kind, ty, m, t, pers=False, wt=False):
    """The Poisson sweep kernel a tune() choice launches (as rocprofv3 names
    it); wt: the weights are looked up by flag pattern (weight mode 1, else
    0: no case of the launch takes the uniform-grid form)."""
    if kind == 0:
        return "k_rbsor<%s, %d, %d>" % (t, ty, m)
    if kind == 1:
        return "k_rbsor_stream<%s, %d, %d>" % (t, m, ty)
    nw = {2: 4, 3: 8, 4: 16}[kind]
    b = lambda x: "true" if x else "false"
    return "k_rbsor_reg<%s, %d, %d, %d, %s, %d>" % (t, m, nw, (ty + 4 * m) // nw, b(pers), 1 if wt else 0)


def reg_updates_per_sweep(NX, NY, kind, ty, m, pers=False):
    """Cell updates one red-black sweep of the register-tile kernel executes
    over an NX x NY interior, averaged over the 2m half-sweeps of a launch:
    every tile computes its whole 128-column, (ty + 4m)-row block (the
    stored region plus the 2m halo it recomputes), except that in a launch
    per stage a wave whose rows all lie within h + 1 rows of the block's top
    or bottom edge skips half-sweep h (no row of it can still be exact)."""
    tw = 128 - 4 * m
    ntiles = -(-NX // tw) * -(-NY // ty)
    nw = {2: 4, 3: 8, 4: 16}[kind]
    rows = (ty + 4 * m) // nw
    live = 0
    for h in range(2 * m):
        for w in range(nw):
            if pers or (w * rows + rows - 1 > h and nw * rows - 1 - w * rows > h):
                live += rows
    return ntiles * 128 * live / (2 * m)


def rel_err(a, b):
    """max |a - b| / max(1, max |b|): the error of a mixed-precision field
    against the FP64 field of the same case and step."""
    import numpy as np
    return float(np.max(np.abs(a - b))) / max(1.0, float(np.max(np.abs(b))))


def fields_err(get_mixed, get_fp64):
    """u, v, p errors; p also modulo the constant of the fluid region
    (serve.pressure_err: the pressure of the all-Neumann problem is fixed
    only up to a constant there, which neither solve pins)."""
    from immerseflow_amd import serve
    out = {f: rel_err(get_mixed(f), get_fp64(f)) for f in ("u", "v", "p")}
    out["p_mod_const"] = serve.pressure_err(get_mixed("p"), get_fp64("p"), get_fp64("flags"))
    return {f: float("%.3g" % v) for f, v in out.items()}


def case_kw(cfg, **over):
    kw = dict(dt=cfg["dt"], re=cfg["re"], bodies=cfg["bodies"], ad_itermax=25, ad_tol=1e-6,
              ppe_itermax=400000, ppe_tol=cfg["ppe_tol"], check_every=160)
    kw.update(over)
    return kw


def new_case(iff, cfg, prec, a, **over):
    c = iff.Case(cfg["xf"], cfg["yf"], precision=prec, **case_kw(cfg, **over))
    c.graphs(not a.no_graphs)
    c.persistent(not a.no_persistent)
    return c


def cpu_threads():
    """Threads for the oracle: OMP_NUM_THREADS when set (the GPU host sets
    it to its per-GPU CPU share), else the cores this process may use."""
    n = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    if n <= 0:
        n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return n


def host_cores():
    """(physical cores, logical CPUs) the host reports (lscpu)."""
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], stdout=subprocess.PIPE, text=True, timeout=10).stdout
        cores = {tuple(ln.split(",")) for ln in out.splitlines() if ln and not ln.startswith("#")}
        return len(cores), os.cpu_count()
    except Exception:
        return None, os.cpu_count()


def oracle_case(cfg, omega, ppe_itermax, check_every):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    prm = dict(dt=cfg["dt"], re=cfg["re"], conv_factor=0.5, ubc=1.0, vbc=0.0, ad_itermax=25, ad_tol=1e-6,
               ppe_solver=1, ppe_omega=omega, ppe_itermax=ppe_itermax, ppe_tol=cfg["ppe_tol"],
               check_every=check_every, ad_omega=1.0)
    bodies = []
    for b in cfg["bodies"]:
        d = dict(b)
        d["c"], d["s"] = b.get("cth", 1.0), b.get("sth", 0.0)
        bodies.append(d)
    return O.PhysCase(cfg["xf"], cfg["yf"], bodies, prm)


def gpu_omega(cfg):
    """The library's default relaxation: 2/(1 + sin(pi/N)), N the larger cell count."""
    n = max(len(cfg["xf"]) - 1, len(cfg["yf"]) - 1)
    return 2.0 / (1.0 + math.sin(math.pi / n))


# ---------------------------------------------------------------------------
# headline: the 4096^2 eight-star workload
# ---------------------------------------------------------------------------
def headline_c3(iff, configs, a):
    cfg = configs.c3()
    NX, NY = len(cfg["xf"]) - 1, len(cfg["yf"]) - 1
    cells = NX * NY
    c = new_case(iff, cfg, iff.IF_MIXED, a)
    st1 = c.step(1)
    log("c3 step 1: sweeps", st1["ppe_sweeps"], "div %.2e" % st1["div_max"])
    t0 = time.perf_counter()
    if a.c3_tune:
        kind, ty, m, k = (int(x) for x in a.c3_tune.split(","))
        c.tune(kind, ty, m, k)
        tr = dict(kernel=kind, tile_y=ty, sweeps_per_launch=m, check_every=k)
    else:
        tr = c.autotune()
    t_tune = time.perf_counter() - t0
    log("c3 autotune", tr, "%.3f s" % t_tune)
    for k in range(max(0, a.warmup - 1)):
        st = c.step(1)
        log("c3 warmup", k + 2, "sweeps", st["ppe_sweeps"], "div %.2e" % st["div_max"])
    c.sweep_timing(reset=True)
    t0 = time.perf_counter()
    st = c.step(a.steps)
    dt = time.perf_counter() - t0
    ms, launches = c.sweep_timing(reset=True)
    sweep_bpc, step_bpc = c.byte_model()
    spl = int(round(sweep_bpc / 13.0))         # sweeps per launch (13 B/cell-sweep, mixed)
    pers = spl > tr["sweeps_per_launch"]
    log("c3 timed %d steps in %.3f s, last sweeps %d, total %d, max div %.2e"
        % (a.steps, dt, st["ppe_sweeps"], st["total_ppe_sweeps"], st["max_div"]))
    avg_s = ms * 1e-3 / max(1, launches)
    upd = int(reg_updates_per_sweep(NX, NY, tr["kernel"], tr["tile_y"], tr["sweeps_per_launch"], pers) * spl)
    useful = cells * spl
    kname = kernel_name(tr["kernel"], tr["tile_y"], tr["sweeps_per_launch"], "float", pers, st["weight_table"])
    uf = bool(st["weight_table"])  # the headline grid is uniform: the five-operation form
    fpu = FLOP_PER_UPDATE[uf]
    achieved = upd * fpu / avg_s / 1e12 if avg_s > 0 else 0.0
    roof = dict(bound="valu", achieved=round(achieved, 2), peak=round(FP32_NONPACKED_TFLOPS, 2), unit="TFLOP/s",
                frac=round(achieved / FP32_NONPACKED_TFLOPS, 4), traffic=None, kernel=kname,
                published_fp32_peak=FP32_PEAK_TFLOPS,
                frac_of_published=round(achieved / FP32_PEAK_TFLOPS, 4),
                updates_per_launch=upd, interior_updates_per_launch=useful,
                recompute_factor=round(upd / useful, 3), sweeps_per_launch=spl,
                flop_per_update=fpu, valu_per_update=VALU_PER_UPDATE[uf],
                mean_launch_us=round(avg_s * 1e6, 3), launches_timed=launches,
                useful_tflops=round(useful * fpu / avg_s / 1e12, 2) if avg_s > 0 else None,
                note="achieved = cell updates the launch executes (the temporal-blocking halo recompute "
                     "included, the halo half-sweeps that no longer hold exact rows left out) x FP32 FLOP per "
                     "update (%s) / mean launch time from device time stamps around each chunk; peak = 256 CUs "
                     "x 64 lanes x 2 FLOP x 2.4 GHz, "
                     "the FP32 rate without packed instructions, which this build does not emit (the "
                     "published 157.3 TFLOP/s counts packed FP32)"
                     % ("9: 1 add + 4 FMA, the form of a grid with one coefficient value per axis" if uf
                        else "11: 1 multiply + 5 FMA"))
    tp = os.path.join(ROOT, "profiles", "traffic_rbsor_reg_c3.json")
    tj = {}
    if os.path.exists(tp):
        with open(tp) as f:  # one entry per profiled shape of the headline kernel
            tj = next((e for e in json.load(f)["kernels"]
                       if e.get("kernel") == kname and e.get("sweeps_per_launch") == spl), {})
        if tj:
            roof["traffic"] = tj.get("bytes_per_launch")
            hb = tj["bytes_per_launch"] / avg_s / 1e9 if avg_s > 0 else None
            roof["hbm"] = dict(achieved=round(hb, 1) if hb else None, peak=HBM_PEAK_GBS, unit="GB/s",
                               frac=round(hb / HBM_PEAK_GBS, 4) if hb else None,
                               source="profiles/traffic_rbsor_reg_c3.json (rocprofv3 FETCH_SIZE, WRITE_SIZE, "
                                      "calibrated) / the HIP-event launch time")
            if tj.get("valu_insts_per_launch"):
                v = tj["valu_insts_per_launch"] * 64 / avg_s / 1e12
                roof["valu_counter"] = dict(lane_ops_per_s_T=round(v, 2),
                                            frac_of_issue=round(v / (FP32_NONPACKED_TFLOPS / 2), 4),
                                            valu_per_update=round(tj["valu_insts_per_launch"] * 64 / upd, 2),
                                            source="SQ_INSTS_VALU (profiles/traffic_rbsor_reg_c3.json) x 64 "
                                                   "lanes / launch time, against 256 CUs x 64 lanes x 2.4 GHz")
    out = dict(case=c, cfg=cfg, tune=tr, cells=cells, NX=NX, NY=NY, seconds=dt, stats=st, roofline=roof,
               step_bpc=step_bpc, mcs=cells * a.steps / dt / 1e6, tune_seconds=t_tune, pers=pers)
    return out


def c3_fp64(iff, cfg, mixed_case, steps, a):
    """The FP64 path on the headline workload (library defaults) for the
    same number of steps from the start: rate and the mixed path's error."""
    c64 = new_case(iff, cfg, iff.IF_FP64, a)
    t0 = time.perf_counter()
    s64 = c64.step(steps)
    dt = time.perf_counter() - t0
    out = dict(mcell_steps_per_s=round((len(cfg["xf"]) - 1) * (len(cfg["yf"]) - 1) * steps / dt / 1e6, 3),
               steps=steps, max_div=s64["max_div"])
    return c64, out


def c3_full(iff, configs, a):
    """--full: all 1000 steps of the headline workload (mixed)."""
    cfg = configs.c3()
    c = new_case(iff, cfg, iff.IF_MIXED, a)
    c.step(1)
    c.autotune()
    t0 = time.perf_counter()
    done = 1
    while done < cfg["steps"]:
        n = min(100, cfg["steps"] - done)
        st = c.step(n)
        done += n
        log("c3 full: %d steps, %.1f s, last sweeps %d, max div %.2e" % (done, time.perf_counter() - t0,
                                                                        st["ppe_sweeps"], st["max_div"]))
    dt = time.perf_counter() - t0
    c.close()
    NX, NY = len(cfg["xf"]) - 1, len(cfg["yf"]) - 1
    return dict(value=round(NX * NY * (cfg["steps"] - 1) / dt / 1e6, 3), unit="Mcell-steps/s",
                steps=cfg["steps"] - 1, seconds=round(dt, 2),
                sample="steps 2..1000 (step 1 untimed, then autotune)")


# ---------------------------------------------------------------------------
# CPU baselines (the oracle)
# ---------------------------------------------------------------------------
def cpu_c3_sample(cfg, gpu_case, gpu_sweeps_per_step, nsw):
    """Oracle steps (FP64, OpenMP) from the GPU's state with the Poisson
    loop capped at nsw and 5*nsw sweeps, each timed twice (the faster run
    kept), after one untimed step that pays the first touch of the oracle's
    arrays: t(step) = t_nsw + (S - nsw) t_sweep, S the GPU's mean sweeps per
    step, t_sweep from the difference."""
    threads = cpu_threads()
    os.environ["OMP_NUM_THREADS"] = str(threads)
    oc = oracle_case(cfg, gpu_omega(cfg), nsw, nsw)
    state = {name: gpu_case.get(name) for name in ("u", "v", "p", "uf", "vf")}
    best = {nsw: math.inf, 5 * nsw: math.inf}
    for i, cap in enumerate((nsw, nsw, 5 * nsw, nsw, 5 * nsw)):
        for name, arr in state.items():
            getattr(oc, name)[:] = arr
        oc.prm.ppe_itermax = cap
        oc.prm.check_every = cap
        t0 = time.perf_counter()
        out = oc.step(1)[-1]
        dt = time.perf_counter() - t0
        assert out[2] == cap
        if i > 0:  # step 0 is the warm-up
            best[cap] = min(best[cap], dt)
    times = [best[nsw], best[5 * nsw]]
    t_sweep = (times[1] - times[0]) / (4 * nsw)
    if not t_sweep > 0:
        raise RuntimeError("CPU sample: the capped steps did not separate (%.3f s, %.3f s)" % tuple(times))
    t_step = times[0] + (gpu_sweeps_per_step - nsw) * t_sweep
    cells = (oc.nx - 2) * (oc.ny - 2)
    phys, logical = host_cores()
    return dict(value=round(cells / t_step / 1e6, 4), unit="Mcell-steps/s", cores=threads, kind="port",
                host_physical_cores=phys, host_logical_cpus=logical,
                sample="headline workload, one FP64 step from the GPU state with the Poisson loop capped at %d "
                       "and %d sweeps, the faster of two runs each after an untimed one (%.2f s, %.2f s); per-step time "
                       "extrapolated to the GPU's %.0f sweeps per step: %.1f s (%.2f ms per sweep)"
                       % (nsw, 5 * nsw, times[0], times[1], gpu_sweeps_per_step, t_step, 1e3 * t_sweep))


def cpu_c1_full(configs):
    cfg = configs.c1()
    threads = cpu_threads()
    os.environ["OMP_NUM_THREADS"] = str(threads)
    oc = oracle_case(cfg, gpu_omega(cfg), 200000, 10)
    t0 = time.perf_counter()
    out = oc.step(cfg["steps"])
    dt = time.perf_counter() - t0
    cells = (oc.nx - 2) * (oc.ny - 2)
    return dict(value=round(cells * cfg["steps"] / dt / 1e6, 4), unit="Mcell-steps/s", cores=threads,
                kind="port", seconds=round(dt, 2), steps=cfg["steps"], max_div=max(o[4] for o in out),
                sweeps_last_step=out[-1][2],
                sample="256^2 cylinder, all 500 FP64 steps from the initial state, stop max|div| <= 1e-8")


# ---------------------------------------------------------------------------
# other workloads
# ---------------------------------------------------------------------------
def gpu_c1(iff, configs, a):
    cfg = configs.c1()
    c = new_case(iff, cfg, iff.IF_FP64, a, check_every=10)
    t0 = time.perf_counter()
    st = c.step(cfg["steps"])
    dt = time.perf_counter() - t0
    c.close()
    return dict(value=round(256 * 256 * cfg["steps"] / dt / 1e6, 3), unit="Mcell-steps/s", steps=cfg["steps"],
                seconds=round(dt, 3), max_div=st["max_div"], sweeps_last_step=st["ppe_sweeps"],
                sample="all 500 steps, FP64, library defaults, stop max|div| <= 1e-8")


def c2_workload(iff, configs, a):
    """The 2048 x 1024 ellipse: all --c2-steps steps mixed (check every 480),
    the FP64 path and the FP32-velocity variant over a window, errors."""
    cfg = configs.c2()
    NX, NY = len(cfg["xf"]) - 1, len(cfg["yf"]) - 1
    cells = NX * NY
    kw = dict(ppe_omega=1.998, ppe_itermax=200000, check_every=480)
    c = new_case(iff, cfg, iff.IF_MIXED, a, **kw)
    c.tune(4, 96, 8, 480)
    t0 = time.perf_counter()
    done, tot = 0, 0
    while done < a.c2_steps:
        n = min(500, a.c2_steps - done)
        st = c.step(n)
        done += n
        tot += st["total_ppe_sweeps"]
        log("c2 %d steps %.2f s, last sweeps %d, max div %.2e" % (done, time.perf_counter() - t0, st["ppe_sweeps"],
                                                                 st["max_div"]))
    dt = time.perf_counter() - t0
    c.close()
    out = dict(value=round(cells * a.c2_steps / dt / 1e6, 3), unit="Mcell-steps/s", steps=a.c2_steps,
               seconds=round(dt, 2), mean_sweeps_per_step=round(tot / a.c2_steps, 1),
               sample="all %d steps from the initial state, mixed, 16 waves x 96 rows x 8 sweeps, persistent, "
                      "check every 480" % a.c2_steps)
    if a.no_fp64:
        return out
    n = a.c2_fp64_steps
    c64 = new_case(iff, cfg, iff.IF_FP64, a, **dict(kw, check_every=160))
    c64.tune(4, 80, 4, 160)
    c64.step(1)
    t0 = time.perf_counter()
    s64 = c64.step(n)
    dt64 = time.perf_counter() - t0
    cm = new_case(iff, cfg, iff.IF_MIXED, a, **kw)
    cm.tune(4, 96, 8, 480)
    t0 = time.perf_counter()
    cm.step(1)
    sm = cm.step(n)
    dtm = time.perf_counter() - t0
    out["fp64"] = dict(value=round(cells * n / dt64 / 1e6, 3), steps="2..%d" % (1 + n), max_div=s64["max_div"],
                       mixed_same_window=round(cells * (1 + n) / dtm / 1e6, 3))
    out["mixed_error_vs_fp64"] = fields_err(cm.get, c64.get)
    cm.close()
    cv = new_case(iff, cfg, iff.IF_MIXED_V32, a, **kw)
    cv.tune(4, 96, 8, 480)
    cv.step(1 + n)
    out["mixed_fp32_velocity_error_vs_fp64"] = fields_err(cv.get, c64.get)
    cv.close()
    c64.close()
    out["error_note"] = ("max|x - FP64| / max(1, max|FP64|) after %d steps from the same start; both solves "
                         "stop at max|div| <= 1e-6" % (1 + n))
    return out


def persistent_tail(cases, sweeps, tr, step):
    """How much of a batch step could run as one persistent launch: the
    running cases' tiles fit on the device (one 16-wave block per CU, 256)
    only once every case with more tiles has converged.  From each case's
    Poisson sweeps in one step (FP64, alone) and its tiles at the batch's
    tuned shape: the sweeps, and the share of the step's tile-sweeps (its
    work), during which the running tiles number at most 256."""
    m = tr["sweeps_per_launch"]
    wo = 128 - 4 * m if tr["kernel"] >= 2 else 128
    tiles = [-(-(len(c["xf"]) - 1) // wo) * -(-(len(c["yf"]) - 1) // tr["tile_y"]) for c in cases]
    cap, smax = 256, max(sweeps)
    work = tail_work = tail_sweeps = 0
    for sw in range(smax):
        act = sum(t for t, n in zip(tiles, sweeps) if n > sw)
        work += act
        if act <= cap:
            tail_work += act
            tail_sweeps += 1
    return dict(step=step, capacity_tiles=cap, tiles_all_cases=sum(tiles), max_sweeps=smax,
                median_sweeps=sorted(sweeps)[len(sweeps) // 2], sweeps_with_tiles_fitting=tail_sweeps,
                work_share_fitting=round(tail_work / max(work, 1), 4))


def run_batch(iff, cases, prec, a, window, tune=None):
    """One batch of the C4 cases in precision prec: step 1 untimed, autotune
    (or the given tune() choice), then steps 2..--batch-steps timed (the
    first `window` of them also on their own).  Returns the rates, the final
    states and the fields of every 16th case after the last step."""
    b = iff.Batch(prec, hbm_budget=0)
    b.graphs(not a.no_graphs)
    b.persistent(not a.no_persistent)
    slots = [b.add(c["xf"], c["yf"], c["steps"], dt=c["dt"], re=c["re"], bodies=c["bodies"],
                   ppe_tol=1e-7, check_every=160) for c in cases]
    b.step(1)
    t0 = time.perf_counter()
    if tune is None:
        tr = b.autotune()
    else:
        b.tune(tune["kernel"], tune["tile_y"], tune["sweeps_per_launch"], tune["check_every"])
        tr = tune
    t_tune = time.perf_counter() - t0
    log("batch tune:", tr, "%.3f s" % t_tune)
    i0 = b.info()
    t0 = time.perf_counter()
    if window:
        b.step(window)
    t_win = time.perf_counter() - t0
    i1 = b.info()
    rest = a.batch_steps - 1 - window
    done = 0
    while done < rest:
        n = min(50, rest - done)
        b.step(n)
        done += n
        log("batch: %d of %d steps, %.1f s" % (1 + window + done, a.batch_steps, time.perf_counter() - t0))
    t_all = time.perf_counter() - t0
    i2 = b.info()
    states = [b.state(s) for s in slots]
    keep = {k: {f: b.get(slots[k], f) for f in ("u", "v", "p")} for k in range(0, len(cases), 16)}
    b.close()
    timed_steps = a.batch_steps - 1
    ia, ib = (i0, i1) if window else (i0, i2)
    ms = ib["sweep_ms"] - ia["sweep_ms"]
    ups = (ib["cell_sweeps"] - ia["cell_sweeps"]) / (ms * 1e-3) if ms > 0 else None
    return dict(cps=len(cases) / (t_all * a.batch_steps / timed_steps), t_all=t_all, t_win=t_win, tr=tr,
                t_tune=t_tune, finished=sum(1 for st in states if st[0] == 1),
                wt=all(st[2]["weight_table"] for st in states),
                max_div=max(st[2]["max_div"] for st in states), keep=keep, ups=ups)


def batched_c4(iff, configs, a):
    """The 256 seeded uneven cases (128^2..1024^2) resident in one batch for
    all --batch-steps steps: step 1 untimed, autotune, steps 2..200 timed;
    the default mixed path (FP32 Poisson error sweeps, FP64 velocity
    passes) and the FP32-stencil path (velocity passes in FP32 too).  Beside
    it: the plain FP64 single-case path (library defaults, one case at a
    time) on every case over steps 2..1+W, and on every 16th case for all
    steps, whose fields after the last step give both batches' errors."""
    cases = configs.c4_cases(256)
    W = a.batch_window
    mx = run_batch(iff, cases, iff.IF_MIXED, a, W)
    # the FP32-stencil batch runs the mixed batch's tuned shape and interval
    v32 = run_batch(iff, cases, iff.IF_MIXED_V32, a, 0, tune=mx["tr"])
    kc = mx["tr"]["check_every"]
    cells = sum((len(c["xf"]) - 1) * (len(c["yf"]) - 1) for c in cases)
    timed_steps = a.batch_steps - 1
    # the FP64 single-case path: every case over the window, every 16th
    # case for all steps (its fields are the reference for the errors)
    tot = 0.0
    last_sweeps = []
    err = {"u": 0.0, "v": 0.0, "p": 0.0, "p_mod_const": 0.0}
    err32 = dict(err)
    for k, c in enumerate(cases):
        s = iff.Case(c["xf"], c["yf"], precision=iff.IF_FP64, dt=c["dt"], re=c["re"], bodies=c["bodies"],
                     ppe_tol=1e-7, check_every=160)
        s.graphs(not a.no_graphs)
        s.persistent(not a.no_persistent)
        s.step(1)
        t0 = time.perf_counter()
        last_sweeps.append(s.step(W)["ppe_sweeps"])
        tot += time.perf_counter() - t0
        s.close()
        if k in mx["keep"]:
            # the error reference: FP64 with the batches' check intervals
            # (160 sweeps in step 1, the tuned one after it), so that every
            # solve stops after the same sweeps and the difference is the
            # precision's.  With other intervals the solves stop at other
            # iterates, each within the tolerance, and since each solve
            # starts from the last pressure, smooth pressure modes keep
            # those differences (p then differs by ~1e-4, u and v by 1e-10)
            r = iff.Case(c["xf"], c["yf"], precision=iff.IF_FP64, dt=c["dt"], re=c["re"], bodies=c["bodies"],
                         ppe_tol=1e-7, check_every=160)
            r.graphs(not a.no_graphs)
            r.step(1)
            r.tune(3, 32, 4, kc)  # any FP64 shape gives the same bits; kc is a multiple of 4
            r.step(a.batch_steps - 1)
            e = fields_err(lambda f: mx["keep"][k][f], r.get)
            err = {f: max(err[f], e[f]) for f in err}
            e = fields_err(lambda f: v32["keep"][k][f], r.get)
            err32 = {f: max(err32[f], e[f]) for f in err32}
            r.close()
        if k % 64 == 63:
            log("batch: FP64 single-case path, %d cases" % (k + 1))
    # the precision's error at a matched residual: four of the cases alone
    # (a case alone gives the same bits as in the batch), mixed, FP32
    # stencils and FP64 with the batch's shape and interval, solved to
    # max|div| <= 1e-10 for 20 steps
    errm = {"u": 0.0, "v": 0.0, "p": 0.0, "p_mod_const": 0.0}
    errm32 = dict(errm)
    for k in (0, 64, 128, 192):
        c = cases[k]
        kw = dict(dt=c["dt"], re=c["re"], bodies=c["bodies"], ppe_tol=1e-10, ppe_itermax=400000, check_every=160)
        runs = {}
        for prec in (iff.IF_MIXED, iff.IF_MIXED_V32, iff.IF_FP64):
            q = iff.Case(c["xf"], c["yf"], precision=prec, **kw)
            q.graphs(not a.no_graphs)
            if prec == iff.IF_FP64:
                q.tune(3, 32, 4, kc)
            else:
                q.tune(mx["tr"]["kernel"], mx["tr"]["tile_y"], mx["tr"]["sweeps_per_launch"], kc)
            q.step(20)
            runs[prec] = {f: q.get(f) for f in ("u", "v", "p", "flags")}
            q.close()
        e = fields_err(lambda f: runs[iff.IF_MIXED][f], lambda f: runs[iff.IF_FP64][f])
        errm = {f: max(errm[f], e[f]) for f in errm}
        e = fields_err(lambda f: runs[iff.IF_MIXED_V32][f], lambda f: runs[iff.IF_FP64][f])
        errm32 = {f: max(errm32[f], e[f]) for f in errm32}
    single_cps_win = len(cases) / (tot * a.batch_steps / W)
    batch_cps_win = len(cases) / (mx["t_win"] * a.batch_steps / W)
    log("batched: %.2f cases/s over steps 2..%d (FP32 stencils %.2f); window %.2f vs FP64 single %.3f cases/s"
        % (mx["cps"], a.batch_steps, v32["cps"], batch_cps_win, single_cps_win))
    tr = mx["tr"]
    rnd = lambda d: {f: float("%.3g" % v) for f, v in d.items()}
    out = dict(value=round(mx["cps"], 3), unit="cases/s", cases=len(cases), steps=a.batch_steps,
               finished=mx["finished"], seconds_steps_2_to_end=round(mx["t_all"], 2),
               mcell_steps_per_s=round(cells * timed_steps / mx["t_all"] / 1e6, 2), max_div=mx["max_div"],
               error_vs_fp64=dict(rnd(err), cases=list(mx["keep"]), after_step=a.batch_steps),
               error_matched_residual=dict(rnd(errm), cases=[0, 64, 128, 192], steps=20, stop="max|div| <= 1e-10"),
               error_note=("error_vs_fp64: the whole run at the workload's stop level; u and v show the precision, "
                           "p also the stop level: max|div| <= 1e-7 leaves the smooth pressure modes undetermined "
                           "to ~1e-4, and a solve that stops one check earlier or later in one precision keeps "
                           "that difference through the following warm-started solves.  error_matched_residual: "
                           "the same cases alone (bit-equal to the batch) solved to 1e-10 in both precisions.  "
                           "p_mod_const: p modulo the constant of the fluid region, the all-Neumann pressure's "
                           "free mode, which the iteration never changes and rounding moves differently in each "
                           "precision"),
               window=dict(steps="2..%d" % (1 + W), batch_cases_per_s=round(batch_cps_win, 3),
                           fp64_single_case_cases_per_s=round(single_cps_win, 4),
                           speedup_vs_fp64_single=round(batch_cps_win / single_cps_win, 2),
                           fp64_cases=len(cases)),
               sweep_gcell_sweeps_per_s=round(mx["ups"] / 1e9, 1) if mx["ups"] else None,
               autotune=dict(kernel=kernel_name(tr["kernel"], tr["tile_y"], tr["sweeps_per_launch"], "float",
                                                wt=mx["wt"]),
                             check_every=tr["check_every"], candidates=tr["candidates"],
                             seconds=round(mx["t_tune"], 3)),
               fp32_stencils=dict(value=round(v32["cps"], 3), unit="cases/s", finished=v32["finished"],
                                  max_div=v32["max_div"],
                                  mcell_steps_per_s=round(cells * timed_steps / v32["t_all"] / 1e6, 2),
                                  error_vs_fp64=dict(rnd(err32), cases=list(v32["keep"]), after_step=a.batch_steps),
                                  error_matched_residual=dict(rnd(errm32), cases=[0, 64, 128, 192], steps=20,
                                                              stop="max|div| <= 1e-10"),
                                  within_mixed_bound=bool(max(errm32["u"], errm32["v"], errm32["p"]) <= 1e-4),
                                  sample="the same batch with the velocity passes in FP32 (IF_MIXED_V32), "
                                         "steps 2..%d timed" % a.batch_steps),
               sample="all 256 cases resident; step 1 untimed, autotune, steps 2..%d timed (value, per case of "
                      "%d steps); window: steps 2..%d of the batch against the FP64 single-case path on all 256 "
                      "cases one after another, same steps, rates per %d-step case; error_vs_fp64: max|x - FP64| "
                      "/ max(1, max|FP64|) after step %d on every 16th case, against the FP64 path run alone "
                      "checking every 160 sweeps in step 1 and every %d after it, as the batches do"
                      % (a.batch_steps, a.batch_steps, 1 + W, a.batch_steps, a.batch_steps, kc))
    out["persistent_tail"] = persistent_tail(cases, last_sweeps, tr, 1 + W)
    return out


def fp64_single_full(iff, configs, a):
    """--full: the FP64 single-case path on all 256 batch cases for all 200
    steps (one case at a time)."""
    cases = configs.c4_cases(256)
    dt = 0.0
    for k, c in enumerate(cases):
        s = iff.Case(c["xf"], c["yf"], precision=iff.IF_FP64, dt=c["dt"], re=c["re"], bodies=c["bodies"],
                     ppe_tol=1e-7, check_every=160)
        s.step(1)
        t0 = time.perf_counter()
        s.step(c["steps"] - 1)
        dt += time.perf_counter() - t0
        s.close()
        if k % 16 == 15:
            log("fp64 single full: %d cases, %.1f s timed" % (k + 1, dt))
    return dict(value=round(len(cases) / (dt * 200 / 199), 4), unit="cases/s", seconds=round(dt, 1),
                sample="all 256 cases, steps 2..200 each (step 1 untimed), one case at a time")


def serving_c5(iff, configs, a):
    """The serving trace's first --serve-requests requests, unclipped (grids
    64^2..2048^2, 100..1000 steps, 20 % heaving cylinders), served at the
    trace's 50 req/s and again with the arrival times stretched so that the
    offered load is half the measured service rate."""
    from immerseflow_amd import serve
    trace = [(t, dict(c, check_every=160)) for t, c in configs.c5_trace(a.serve_requests)]
    budget = 64 << 30
    small = sorted(range(len(trace)), key=lambda k: (len(trace[k][1]["xf"]) ** 2) * trace[k][1]["steps"])[:2]
    recs, summ = serve.serve(trace, precision=iff.IF_MIXED, hbm_budget=budget, tune="auto", keep_fields=set(small))
    log("serving (50 req/s):", summ)
    # sub-saturating: mean inter-arrival = 2x the mean service time per request
    svc = summ["wall_s"] / max(1, summ["completed"])
    span = trace[-1][0] if trace[-1][0] > 0 else 1.0
    scale = 2.0 * svc * len(trace) / span
    recs2, summ2 = serve.serve(trace, precision=iff.IF_MIXED, hbm_budget=budget, tune="auto", time_scale=scale)
    log("serving (sub-saturating):", summ2)
    err = {"u": 0.0, "v": 0.0, "p": 0.0, "p_mod_const": 0.0}
    for r in recs:
        if r["id"] in small and r["state"] == 1:
            e = serve.error_vs_fp64(r, trace[r["id"]][1])
            err = {f: max(err[f], e[f]) for f in err}

    def pick(s):
        return dict(cases_per_s=round(s["cases_per_s"], 4), p50_s=round(s["p50_s"], 3), p99_s=round(s["p99_s"], 3),
                    mcell_steps_per_s=round(s["mcell_steps_per_s"], 2), max_div=s["max_div"],
                    max_resident=s["max_resident"], wall_s=round(s["wall_s"], 2), completed=s["completed"],
                    autotunes=s["autotunes"], last_autotune=s["autotune"])
    sizes = [len(c["xf"]) - 1 for _, c in trace]
    return dict(value=round(summ["cases_per_s"], 4), unit="cases/s", requests=len(trace),
                at_50_req_per_s=pick(summ), sub_saturating=dict(pick(summ2), arrival_scale=round(scale, 2),
                                                                 offered_req_per_s=round(len(trace) / (span * scale), 4)),
                budget_bytes=budget, grids=sizes, steps=[c["steps"] for _, c in trace],
                heaving=sum(1 for _, c in trace if c["bodies"][0].get("motion", 0)),
                error_vs_fp64=dict({f: float("%.3g" % v) for f, v in err.items()}, requests=small),
                sample="first %d requests of the seeded trace, unclipped (n %d..%d, %d..%d steps), served at 50 "
                       "req/s and with arrivals stretched to half the measured service rate; error_vs_fp64: the "
                       "two smallest requests' served fields against the FP64 path alone"
                       % (len(trace), min(sizes), max(sizes), min(c["steps"] for _, c in trace),
                          max(c["steps"] for _, c in trace)))


def line_sor_f2(iff, configs, a):
    """The deck's PPE_Solver 2 (zebra line SOR): whole steps of the 256^2
    cylinder to its 1e-8 stop, FP64 and mixed; and the time per sweep on the
    2048x1024 workload, from the sweep events of a solve capped at 400
    sweeps."""
    out = {}
    want = set(a.line_cases.split(","))
    for cname, mk, steps, cap in (("c1", configs.c1, 2, 200000), ("c2", configs.c2, 1, 400)):
        if cname not in want:
            continue
        cfg = mk()
        NX, NY = len(cfg["xf"]) - 1, len(cfg["yf"]) - 1
        for pname, prec in (("fp64", iff.IF_FP64), ("mixed", iff.IF_MIXED)):
            c = iff.Case(cfg["xf"], cfg["yf"], precision=prec, dt=cfg["dt"], re=cfg["re"], bodies=cfg["bodies"],
                         ad_itermax=25, ad_tol=1e-6, ppe_solver=iff.IF_PPE_LINE, ppe_itermax=cap,
                         ppe_tol=cfg["ppe_tol"], check_every=40)
            c.graphs(not a.no_graphs)
            c.step(1)
            c.sweep_timing(reset=True)
            t0 = time.perf_counter()
            st = c.step(steps)
            dt = time.perf_counter() - t0
            ms, nsw = c.sweep_timing(reset=True)
            c.close()
            us = 1e3 * ms / max(1, nsw)
            r = dict(us_per_sweep=round(us, 2), mcell_sweeps_per_s=round(NX * NY / us, 1) if us > 0 else None)
            if cname == "c1":
                r.update(ms_per_step=round(1e3 * dt / steps, 2), sweeps_last_step=st["ppe_sweeps"],
                         max_div=st["max_div"])
            out["%s_%s" % (cname, pname)] = r
            log("line sor", cname, pname, r)
    out["sample"] = ("c1: 256^2 cylinder, step 1 untimed, steps 2..3 timed to max|div| <= 1e-8; c2: 2048x1024, "
                     "one step with the solve capped at 400 sweeps; us_per_sweep from HIP events around the "
                     "sweep chunks (both colours)")
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    import immerseflow_amd as iff
    from immerseflow_amd import configs
    n, arch, cus, mem = iff.device_query()
    log("device", arch, cus, "CUs")

    h = headline_c3(iff, configs, a)
    case = h["case"]
    cpu = None
    cpu_extra = {}
    if not a.no_cpu:
        S = h["stats"]["total_ppe_sweeps"] / a.steps
        cpu = cpu_c3_sample(h["cfg"], case, S, a.cpu_sweeps)
        log("cpu baseline (headline sample)", cpu)
    fp64 = None
    if not a.no_fp64:
        nst = 1 + max(0, a.warmup - 1) + a.steps
        n64 = 3
        c64, fp64 = c3_fp64(iff, h["cfg"], case, n64, a)
        cm = new_case(iff, h["cfg"], iff.IF_MIXED, a)
        cm.step(1)
        cm.autotune()
        cm.step(n64 - 1)
        fp64["mixed_error_vs_fp64"] = fields_err(cm.get, c64.get)
        fp64["error_note"] = ("max|mixed - FP64| / max(1, max|FP64|) after %d steps from the same start, both "
                              "stopping at max|div| <= 1e-7; FP64 with library defaults" % n64)
        cm.close()
        c64.close()
        del nst
    case.close()
    full = c3_full(iff, configs, a) if a.full else None

    c2 = None if a.no_c2 else c2_workload(iff, configs, a)
    c1 = None
    if not a.no_c1:
        c1 = dict(gpu_fp64=gpu_c1(iff, configs, a))
        if not a.no_cpu:
            c1["cpu_oracle"] = cpu_c1_full(configs)
            c1["gpu_vs_cpu"] = round(c1["gpu_fp64"]["value"] / c1["cpu_oracle"]["value"], 2)
            cpu_extra["c1_full"] = c1["cpu_oracle"]
        log("c1", c1)
    batched = None if a.no_batch else batched_c4(iff, configs, a)
    if batched is not None and a.full:
        batched["fp64_single_case_full"] = fp64_single_full(iff, configs, a)
        batched["speedup_vs_fp64_single_full"] = round(batched["value"] /
                                                       batched["fp64_single_case_full"]["value"], 2)
    serving = None if a.no_serve else serving_c5(iff, configs, a)
    line = None if a.no_line else line_sor_f2(iff, configs, a)

    st = h["stats"]
    tr = h["tune"]
    out = {
        "metric": "Mcell-steps/s + % HBM roofline; batched cases/s and p50/p99 latency at div<=1e-6",
        "value": round(h["mcs"], 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * h["seconds"] / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "C3 eight seeded star bodies, 4096x4096 uniform, 20x20 box",
                   "NX": h["NX"], "NY": h["NY"], "dt": h["cfg"]["dt"], "Re": h["cfg"]["re"],
                   "bodies": len(h["cfg"]["bodies"]),
                   "precision": "mixed (FP32 Poisson error sweeps in registers, FP64 residual/defect "
                                "correction, FP64 velocity passes)",
                   "poisson": "red-black SOR, stop max|div|<=%g" % h["cfg"]["ppe_tol"],
                   "timed_steps": "%d..%d of %d" % (a.warmup + 1, a.warmup + a.steps, h["cfg"]["steps"]),
                   "sweeps_last_step": st["ppe_sweeps"], "mean_sweeps_per_step": st["total_ppe_sweeps"] / a.steps,
                   "ad_iters_last_step": st["ad_iters"], "max_div": st["max_div"], "parallelism": "single GPU",
                   "autotune": dict(kernel=h["roofline"]["kernel"], check_every=tr["check_every"],
                                    seconds=round(h["tune_seconds"], 3)),
                   "poisson_launch": ("persistent launch per check interval" if h["pers"] else
                                      "launch per %d sweeps" % tr["sweeps_per_launch"]) +
                                     ("" if a.no_graphs else ", hipGraph replay")},
        "roofline": h["roofline"],
        "cpu_baseline": cpu,
        "cpu_baselines": cpu_extra or None,
        "fp64": fp64,
        "full_length": full,
        "ellipse_2048x1024": c2,
        "cylinder_256": c1,
        "batched": batched,
        "serving": serving,
        "line_sor": line,
        "step_bytes_per_cell": round(h["step_bpc"], 1),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
