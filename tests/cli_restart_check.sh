#!/bin/bash
# The immerseflow CLI restart path on one GPU: 6 steps straight vs 4 steps,
# restart from the lossless checkpoint, 2 more; the final checkpoints must be
# byte-identical.  Usage: tests/cli_restart_check.sh [work dir, default a new temporary one]  (needs a GPU
# and the built CLI, immerseflow_amd/immerseflow)
set -e
cd "$(dirname "$0")/.."
W=${1:-$(mktemp -d)}; mkdir -p $W/in $W/r1 $W/r2 $W/r3 $W/ckA $W/ckB $W/ckC
W=$W python3 - <<'PY'
import numpy as np
import os; W = os.environ["W"]
def grid(path, f):
    with open(path, "w") as fp:
        for i, v in enumerate(f): fp.write("%10d %.16E\n" % (i + 1, v))
grid(W + "/in/xgrid.dat2", np.linspace(0, 8, 97)); grid(W + "/in/ygrid.dat2", np.linspace(0, 4, 49))
open(W + "/in/inputs.txt", "w").write("""Restart     Restart_Time
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
""")
PY
B=immerseflow_amd/immerseflow
A="--deck $W/in/inputs.txt --xgrid $W/in/xgrid.dat2 --ygrid $W/in/ygrid.dat2"
timeout -k 10 120 $B $A --results $W/r1 --steps 6 --checkpoint $W/ckA
timeout -k 10 120 $B $A --results $W/r2 --steps 4 --checkpoint $W/ckB
timeout -k 10 120 $B $A --results $W/r3 --steps 2 --restart $W/ckB --checkpoint $W/ckC
for f in uc.dat vc.dat p.dat uf.dat vf.dat; do cmp $W/ckA/$f $W/ckC/$f && echo "same $f"; done
