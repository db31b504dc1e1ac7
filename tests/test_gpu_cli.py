"""GPU: the immerseflow CLI, the drop-in for the reference's main()
(src/main.cu:75-102), run as a fresh child process.

* With no arguments it reads ../inputs/inputs.txt (main.cu:80) and the grid
  files ../inputs/xgrid.dat2, ygrid.dat2 (preSim.cu:268,281), writes
  ../results/final_results.dat at initialisation (preSim.cu:217) and
  uc.dat, vc.dat after every ADsolver call (ADSolver.cu:378-379).  On the
  reference's own deck and grids, run for the 20 steps the fixtures hold,
  the three files equal the reference's results byte for byte.
* Restart (the deck's Restart flag, main.cu:27-30): 6 steps straight and
  4 steps + restart from the lossless checkpoint + 2 steps end in
  byte-identical checkpoints.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "immerseflow_amd", "immerseflow")


def _run(args, cwd, timeout=120):
    print("[gpu-test] cli", " ".join(args), "in", cwd, flush=True)
    r = subprocess.run([CLI] + args, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=timeout)
    assert r.returncode == 0, r.stdout[-2000:]
    return r.stdout


def test_cli_replays_reference_main_on_its_deck(require_gpu, tmp_path):
    assert os.path.exists(CLI), "build the CLI first (python -m immerseflow_amd.build)"
    for d in ("src", "inputs", "results"):
        (tmp_path / d).mkdir()
    deck = open(os.path.join(GOLDEN, "inputs.txt")).read()
    # the fixtures are the state after 20 ADsolver calls (dt 1e-3): tmax 20
    lines = deck.splitlines()
    k = next(i for i, ln in enumerate(lines) if "ErrorMax" in ln)
    vals = lines[k + 1].split()
    assert vals[1] == "100"
    vals[1] = "20"
    lines[k + 1] = "    ".join(vals)
    (tmp_path / "inputs" / "inputs.txt").write_text("\n".join(lines) + "\n")
    for g in ("xgrid.dat2", "ygrid.dat2"):
        shutil.copy(os.path.join(GOLDEN, g), tmp_path / "inputs" / g)
    out = _run([], cwd=str(tmp_path / "src"))
    assert out.count("5 AD iterations") == 20
    for name in ("uc.dat", "vc.dat", "final_results.dat"):
        got = (tmp_path / "results" / name).read_bytes()
        assert got == open(os.path.join(GOLDEN, name), "rb").read(), name


def _grid(path, f):
    with open(path, "w") as fp:
        for i, v in enumerate(f):
            fp.write("%10d %.16E\n" % (i + 1, v))


DECK = """Restart     Restart_Time
0           0
nx      ny
96      48
Lx      Ly
8      4
w-AD    w-PPE   AD-itermax  PPE-itermax  AD_Solver PPE_Solver(1. Point GS, 2. Line SOR)
1       1       25          100000          1         1
ErrorMax    tmax    dt       Re      mu
1E-6        6       0.01    100    0.01
Write Interval(t/dt)
100
___________________________| Fractional Step |_______________________________
Mode   Precision   Conv-factor   Omega-PPE   Tol-PPE   Check-every
physics  fp64   0.5   1.8   1e-9   10
___________________________| Bodies |_______________________________
Body (circle xc yc r)
circle 3 2 0.5
"""


def test_cli_restart_continues_bit_for_bit(require_gpu, tmp_path):
    W = tmp_path
    for d in ("in", "r1", "r2", "r3", "ckA", "ckB", "ckC"):
        (W / d).mkdir()
    _grid(W / "in" / "xgrid.dat2", np.linspace(0, 8, 97))
    _grid(W / "in" / "ygrid.dat2", np.linspace(0, 4, 49))
    (W / "in" / "inputs.txt").write_text(DECK)
    A = ["--deck", str(W / "in" / "inputs.txt"), "--xgrid", str(W / "in" / "xgrid.dat2"),
         "--ygrid", str(W / "in" / "ygrid.dat2")]
    _run(A + ["--results", str(W / "r1"), "--steps", "6", "--checkpoint", str(W / "ckA")], cwd=str(W))
    _run(A + ["--results", str(W / "r2"), "--steps", "4", "--checkpoint", str(W / "ckB")], cwd=str(W))
    out = _run(A + ["--results", str(W / "r3"), "--steps", "2", "--restart", str(W / "ckB"),
                    "--checkpoint", str(W / "ckC")], cwd=str(W))
    assert "with faces" in out
    for f in ("uc.dat", "vc.dat", "p.dat", "uf.dat", "vf.dat"):
        assert (W / "ckA" / f).read_bytes() == (W / "ckC" / f).read_bytes(), f
    # the reference-format outputs of the continued run equal the straight run's
    for f in ("uc.dat", "vc.dat", "p.dat"):
        assert (W / "r1" / f).read_bytes() == (W / "r3" / f).read_bytes(), f
