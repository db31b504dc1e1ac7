"""GPU: restart (the deck's Restart flag; the predecessor's restart from
data files).  A checkpoint written at 17 digits (u, v, p and the faces)
restarts a case exactly: the continued run equals the uninterrupted one bit
for bit.  The reference's own %f outputs (u, v, p only) restart a case whose
faces are interpolated from u and v, which continues to the stop level."""
import numpy as np
import pytest

from test_gpu_parity import CASES, rel

import immerseflow_amd as iff

pytestmark = pytest.mark.gpu


def _kw(prec):
    xf, yf, bodies = CASES["ellipse"]
    return xf, yf, dict(precision=prec, bodies=bodies, dt=0.01, re=100.0, ppe_omega=1.8, ppe_itermax=200000,
                        ppe_tol=1e-9, check_every=10)


@pytest.mark.parametrize("prec", [iff.IF_FP64, iff.IF_MIXED])
def test_checkpoint_restart_is_exact(require_gpu, prec, tmp_path):
    print("[gpu-test] checkpoint restart", prec, flush=True)
    xf, yf, kw = _kw(prec)
    a = iff.Case(xf, yf, **kw)
    a.step(3)
    a.write_checkpoint(str(tmp_path))
    b = iff.Case(xf, yf, **kw)
    b.restart_from(str(tmp_path))
    for f in ("u", "v", "p", "uf", "vf"):
        assert np.array_equal(a.get(f), b.get(f)), f
    for _ in range(2):
        sa, sb = a.step(1), b.step(1)
        assert sa["ad_iters"] == sb["ad_iters"] and sa["ppe_sweeps"] == sb["ppe_sweeps"]
    for f in ("u", "v", "p", "uf", "vf"):
        assert np.array_equal(a.get(f), b.get(f)), f
    a.close()
    b.close()


def test_restart_from_reference_outputs(require_gpu, tmp_path):
    """u, v, p alone (faces interpolated): from the reference's %f files the
    run continues as from the same fields at full precision, to the files'
    rounding, and reaches the stop level."""
    print("[gpu-test] restart from %f files", flush=True)
    xf, yf, kw = _kw(iff.IF_FP64)
    a = iff.Case(xf, yf, **kw)
    a.step(3)
    (tmp_path / "ref").mkdir()
    (tmp_path / "full").mkdir()
    for name, field in (("uc.dat", "u"), ("vc.dat", "v"), ("p.dat", "p")):
        iff.write_tecplot(str(tmp_path / "ref" / name), a.xc, a.yc, a.get(field))
        iff.write_tecplot(str(tmp_path / "full" / name), a.xc, a.yc, a.get(field), digits=17)
    b, c = iff.Case(xf, yf, **kw), iff.Case(xf, yf, **kw)
    b.restart_from(str(tmp_path / "ref"))
    c.restart_from(str(tmp_path / "full"))
    assert np.array_equal(c.get("u"), a.get("u")) and rel(b.get("u"), a.get("u")) <= 1e-6
    sb, sc = b.step(2), c.step(2)
    for st in (sb, sc):
        assert st["max_div"] <= kw["ppe_tol"] and st["failed"] == 0
    for f in ("u", "v", "p"):
        assert rel(b.get(f), c.get(f)) < 1e-4, f
    for x in (a, b, c):
        x.close()


def _read_f(path, n):
    """Values of a file in the writer's format, parsed here independently of
    the library's reader (header of 3 lines, then 'x,y,value' rows)."""
    v = np.loadtxt(path, skiprows=3, delimiter=",", usecols=2)
    assert v.size == n
    return np.ascontiguousarray(v)


@pytest.mark.parametrize("name", ["ellipse", "circle-stretched"])
def test_restart_from_files_matches_oracle(require_gpu, name, tmp_path):
    """The GPU and the oracle restart from the same %f files (u, v, p; faces
    interpolated) and then agree step for step: sweep counts identical,
    fields to 1e-10 (the restart path of the deck's Restart flag,
    src/main.cu:27-30; the predecessor's, test/UTIL_PRE_SIM.f90:130-164)."""
    import oracle_lib as O
    print("[gpu-test] restart from %f files vs oracle", name, flush=True)
    xf, yf, bodies = CASES[name]
    prm = dict(dt=0.01, re=100.0, conv_factor=0.5, ubc=1.0, vbc=0.0, ad_itermax=25, ad_tol=1e-6, ppe_solver=1,
               ppe_omega=1.8, ppe_itermax=200000, ppe_tol=1e-9, check_every=10)
    a = iff.Case(xf, yf, bodies=bodies, **prm)
    a.step(3)
    for fname, field in (("uc.dat", "u"), ("vc.dat", "v"), ("p.dat", "p")):
        iff.write_tecplot(str(tmp_path / fname), a.xc, a.yc, a.get(field))
    a.close()
    n = len(a.xc) * len(a.yc)
    u, v, p = (_read_f(str(tmp_path / f), n) for f in ("uc.dat", "vc.dat", "p.dat"))
    g = iff.Case(xf, yf, bodies=bodies, **prm)
    g.restart_from(str(tmp_path))
    ob = []
    for b in bodies:
        d = dict(b)
        d["c"], d["s"] = b.get("cth", 1.0), b.get("sth", 0.0)
        ob.append(d)
    o = O.PhysCase(xf, yf, ob, prm)
    o.restart(u, v, p)
    for f, ref in (("u", o.u), ("v", o.v), ("p", o.p), ("uf", o.uf), ("vf", o.vf)):
        assert rel(g.get(f), ref) <= 1e-13, f
    for k in range(3):
        sg = g.step(1)
        so = o.step(1)[-1]
        assert (sg["ad_iters"], sg["ppe_sweeps"]) == (so[0], so[2]), k
    for f, ref in (("u", o.u), ("v", o.v), ("p", o.p), ("uf", o.uf), ("vf", o.vf)):
        assert rel(g.get(f), ref) <= 1e-10, f
    g.close()


# ---------------------------------------------------------------------------
# monitors: probes and body force
# ---------------------------------------------------------------------------
def _force_numpy(xc, yc, dx, dy, flags, u, v, p, mu):
    """The monitor's force rule restated on the host from the fields."""
    nx, ny = len(xc), len(yc)
    F = flags.reshape(ny, nx)
    U, V, Pp = u.reshape(ny, nx), v.reshape(ny, nx), p.reshape(ny, nx)
    fluid = (F & 1) != 0
    fx = fy = 0.0
    for j in range(1, ny - 1):
        for i in range(1, nx - 1):
            if not fluid[j, i]:
                continue
            if not fluid[j, i + 1]:
                fx += Pp[j, i] * dy[j]
                fy += mu * V[j, i] * 2.0 / dx[i] * dy[j]
            if not fluid[j, i - 1]:
                fx -= Pp[j, i] * dy[j]
                fy += mu * V[j, i] * 2.0 / dx[i] * dy[j]
            if not fluid[j + 1, i]:
                fy += Pp[j, i] * dx[i]
                fx += mu * U[j, i] * 2.0 / dy[j] * dx[i]
            if not fluid[j - 1, i]:
                fy -= Pp[j, i] * dx[i]
                fx += mu * U[j, i] * 2.0 / dy[j] * dx[i]
    return fx, fy


def test_monitor_probes_and_force(require_gpu):
    print("[gpu-test] monitors", flush=True)
    xf, yf, kw = _kw(iff.IF_FP64)
    c = iff.Case(xf, yf, **kw)
    probes = [(2.5, 1.0), (0.3, 1.7), (3.9, 0.05)]
    c.monitor(probes, capacity=8)
    c.step(3)
    rec = c.monitor_read()
    assert rec.shape == (3, 3 + 3 * len(probes))
    assert np.allclose(rec[:, 0], kw["dt"] * np.arange(1, 4), rtol=0, atol=1e-15)
    u, v, p = c.get("u"), c.get("v"), c.get("p")
    nx = len(c.xc)
    for k, (x, y) in enumerate(probes):
        i = 1 + int(np.argmin(np.abs(c.xc[1:-1] - x)))
        j = 1 + int(np.argmin(np.abs(c.yc[1:-1] - y)))
        assert tuple(rec[-1, 3 + 3 * k:6 + 3 * k]) == (u[i + j * nx], v[i + j * nx], p[i + j * nx])
    fx, fy = _force_numpy(c.xc, c.yc, c.dx, c.dy, c.get("flags").astype(np.uint8), u, v, p, 1.0 / kw["re"])
    assert abs(rec[-1, 1] - fx) <= 1e-12 * abs(fx) and abs(rec[-1, 2] - fy) <= 1e-12 * max(abs(fy), abs(fx))
    assert fx > 0  # drag along the free stream
    assert c.monitor_read().shape[0] == 0  # a read empties the ring
    c.step(10)  # more steps than the ring holds: the newest 8 are kept
    rec = c.monitor_read()
    assert rec.shape[0] == 8 and np.allclose(rec[:, 0], kw["dt"] * np.arange(6, 14), rtol=0, atol=1e-14)
    c.close()
