"""Physics check of the completed fractional step (needs a GPU; about a
minute): flow past a circular cylinder, D = 1, at Re = 100 on the 16 x 16,
256^2 grid of the cylinder workload, run until the wake sheds.  Reports the
mean drag coefficient and the Strouhal number from the lift signal recorded
by the force monitor.  The cylinder sits a sixteenth of a cell above the
domain's mid-line so the symmetric start breaks within the run.  The
convective term is taken at full weight (conv_factor 1): the reference's
ADSource halves it (ADSolver.cu:66-74), which is kept as the solver's
default for parity but is not the Navier-Stokes equations.

    python tests/cylinder_shedding_check.py [--t-end 200] [--dt 0.01] [--conv-factor 1]
"""
import argparse
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import immerseflow_amd as iff  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--t-end", type=float, default=200.0)
    ap.add_argument("--conv-factor", type=float, default=1.0)
    ap.add_argument("--dt", type=float, default=0.01)
    ap.add_argument("--precision", default="mixed")
    a = ap.parse_args()
    xf = np.linspace(0.0, 16.0, 257)
    prec = iff.IF_MIXED if a.precision == "mixed" else iff.IF_FP64
    c = iff.Case(xf, xf, precision=prec, bodies=[iff.circle(5.0, 8.0 + 1.0 / 256, 0.5)], dt=a.dt, re=100.0,
                 conv_factor=a.conv_factor,
                 ppe_tol=1e-6, ppe_itermax=200000, check_every=40)
    c.monitor([(7.0, 8.0)], capacity=1 << 20)
    nsteps = int(round(a.t_end / a.dt))
    done = 0
    while done < nsteps:
        n = min(1000, nsteps - done)
        st = c.step(n)
        done += n
        print("t = %.1f  sweeps %d  max|div| %.2e" % (done * a.dt, st["ppe_sweeps"], st["max_div"]), flush=True)
    rec = c.monitor_read()
    t, cd, cl = rec[:, 0], 2.0 * rec[:, 1], 2.0 * rec[:, 2]
    late = t >= t[-1] - 40.0
    s = cl[late] - cl[late].mean()
    up = np.nonzero((s[:-1] < 0) & (s[1:] >= 0))[0]
    tz = t[late][up] - s[up] * (t[late][up + 1] - t[late][up]) / (s[up + 1] - s[up])
    st_num = (len(tz) - 1) / (tz[-1] - tz[0]) if len(tz) > 2 else float("nan")
    print("mean Cd %.3f  Cl amplitude %.3f  Strouhal %.4f  (last 40 time units, %d lift periods)"
          % (cd[late].mean(), 0.5 * (cl[late].max() - cl[late].min()), st_num, max(0, len(tz) - 1)))
    c.close()


if __name__ == "__main__":
    main()
