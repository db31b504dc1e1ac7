"""The oracle against the reference's own output files (CPU).

results/uc.dat, vc.dat, p.dat and final_results.dat were written by the
reference solver; the oracle must reproduce every printed digit.
"""
import ctypes as C
import os

import numpy as np

import oracle_lib as O
from conftest import GOLDEN


def _fixture_grid():
    xf = O.read_grid_file(os.path.join(GOLDEN, "xgrid.dat2"))
    yf = O.read_grid_file(os.path.join(GOLDEN, "ygrid.dat2"))
    xc, dx = O.metric(xf)
    yc, dy = O.metric(yf)
    return xc, dx, yc, dy


def _read(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return f.read()


def test_grid_metric_matches_fixture_coordinates():
    xc, dx, yc, dy = _fixture_grid()
    assert len(xc) == 52 and len(yc) == 52
    rows = _read("final_results.dat").splitlines()[3:]
    xs = [float(r.split(",")[0]) for r in rows[:52]]
    ys = [float(r.split(",")[1]) for r in rows[::52]]
    assert ["%f" % v for v in xc] == ["%f" % v for v in xs]
    assert ["%f" % v for v in yc] == ["%f" % v for v in ys]
    # ghost mirrors: west about x = 0, east about the last face
    assert xc[0] == -xc[1] and abs(xc[-1] - (2.0 * 1.0 - xc[-2])) < 1e-15
    assert dx[0] == dx[1] and dx[-1] == dx[-2]


def test_iblank_stub_fixture():
    xc, _, yc, _ = _fixture_grid()
    ib = np.ones(52 * 52)
    assert O.format_tecplot(52, 52, xc, yc, ib) == _read("final_results.dat")


def _compat_run(steps):
    L = O.lib()
    xc, dx, yc, dy = _fixture_grid()
    nx, ny = len(xc), len(yc)
    u, v, p = np.zeros(nx * ny), np.zeros(nx * ny), np.zeros(nx * ny)
    uf, vf = np.zeros((nx - 1) * (ny - 2)), np.zeros((nx - 2) * (ny - 1))
    L.of_compat_init(nx, ny, O.dptr(xc), O.dptr(yc), O.dptr(u), O.dptr(v), O.dptr(uf),
                     O.dptr(vf), O.dptr(p))
    ib = np.ones(nx * ny)
    iters = []
    for _ in range(steps):
        r = C.c_double()
        iters.append(L.of_compat_ad_step(nx, ny, O.dptr(dx), O.dptr(dy), 0.001, 150.0, 25,
                                         O.dptr(ib), O.dptr(u), O.dptr(v), O.dptr(uf),
                                         O.dptr(vf), C.byref(r)))
    return xc, yc, u, v, iters


def test_compat_adsolver_reproduces_uc_vc():
    xc, yc, u, v, iters = _compat_run(20)
    assert iters == [5] * 20
    assert O.format_tecplot(52, 52, xc, yc, u) == _read("uc.dat")
    assert O.format_tecplot(52, 52, xc, yc, v) == _read("vc.dat")


def test_compat_adsolver_step_count_is_pinned():
    # 19 and 21 steps do not reproduce the fixture: the match pins the count
    xc, yc, u, v, _ = _compat_run(19)
    assert O.format_tecplot(52, 52, xc, yc, u) != _read("uc.dat")


def test_compat_ppe_reproduces_p():
    L = O.lib()
    xc, dx, yc, dy = _fixture_grid()
    nx, ny = len(xc), len(yc)
    p = np.zeros(nx * ny)
    r = C.c_double()
    sweeps = L.of_compat_ppe(nx, ny, O.dptr(dx), O.dptr(dy), 100000, O.dptr(p), C.byref(r))
    # the signed residual is summed in the order of the reference's Reduction
    # (reduce6<256>, preSim.cu:22-50,376-445); the fixture itself is matched
    # from sweep 12,606 on, so it does not fix the count
    assert sweeps == 14807 and r.value <= 1e-6
    assert O.format_tecplot(nx, ny, xc, yc, p) == _read("p.dat")


def test_compat_ppe_equals_direct_laplace_solve():
    """p.dat's closed form: the 5-point Laplace problem with Dirichlet ghost
    values 100 (W, S) and 0 (E, N), solved directly."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    n = 50
    N = n * n
    A = sp.lil_matrix((N, N))
    b = np.zeros(N)
    for j in range(n):
        for i in range(n):
            k = i + j * n
            A[k, k] = -4.0
            for di, dj, bc in ((1, 0, 0.0), (-1, 0, 100.0), (0, 1, 0.0), (0, -1, 100.0)):
                ii, jj = i + di, j + dj
                if 0 <= ii < n and 0 <= jj < n:
                    A[k, ii + jj * n] = 1.0
                else:
                    b[k] -= bc
    sol = spl.spsolve(A.tocsr(), b)
    rows = _read("p.dat").splitlines()[3:]
    vals = np.array([float(r.split(",")[2]) for r in rows]).reshape(52, 52)
    assert np.max(np.abs(vals[1:51, 1:51] - sol.reshape(n, n))) < 1e-6
