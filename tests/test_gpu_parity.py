"""GPU parity: the HIP path (through the C ABI) against the oracle and the
reference's golden files.

Tolerances: integer and index outputs (flags, ghost lists, sweep counts)
must be identical; FP64 fields agree to 1e-10 relative (the kernels
evaluate the oracle's expressions in the same order without contraction, so
they are normally bit-identical; only reduction order differs); the mixed
path (FP32 stencils and Poisson sweeps with FP64 refinement) agrees with the
FP64 oracle to 1e-4 relative at the same Poisson stop level.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as O
from conftest import GOLDEN

import immerseflow_amd as iff

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-test]", *a, flush=True)


def rel(a, b):
    s = max(1.0, float(np.max(np.abs(b))))
    return float(np.max(np.abs(a - b))) / s


def stretched_faces(n, L, ratio):
    w = ratio ** np.arange(n)
    w = w / w.sum() * L
    return np.concatenate([[0.0], np.cumsum(w)])


# ---------------------------------------------------------------------------
# compat mode: the reference's own outputs
# ---------------------------------------------------------------------------
def test_compat_adsolver_golden(require_gpu, tmp_path):
    _say("compat ADsolver 52x52 x 20 steps")
    xf = O.read_grid_file(os.path.join(GOLDEN, "xgrid.dat2"))
    c = iff.Case(xf, xf, mode=iff.IF_MODE_COMPAT, dt=0.001, re=150.0, ad_itermax=25)
    iters = [c.ad_step()["ad_iters"] for _ in range(20)]
    assert iters == [5] * 20
    iff.write_tecplot(str(tmp_path / "uc.dat"), c.xc, c.yc, c.get("u"))
    iff.write_tecplot(str(tmp_path / "vc.dat"), c.xc, c.yc, c.get("v"))
    for name in ("uc.dat", "vc.dat"):
        assert (tmp_path / name).read_bytes() == open(os.path.join(GOLDEN, name), "rb").read(), name
    # against the oracle's fields (the IC uses the device's exp/pow)
    xc, dx = O.metric(xf)
    n = len(xc)
    u, v, p = np.zeros(n * n), np.zeros(n * n), np.zeros(n * n)
    uf, vf = np.zeros((n - 1) * (n - 2)), np.zeros((n - 2) * (n - 1))
    L = O.lib()
    L.of_compat_init(n, n, O.dptr(xc), O.dptr(xc), O.dptr(u), O.dptr(v), O.dptr(uf), O.dptr(vf), O.dptr(p))
    ib = np.ones(n * n)
    for _ in range(20):
        r = C.c_double()
        L.of_compat_ad_step(n, n, O.dptr(dx), O.dptr(dx), 0.001, 150.0, 25, O.dptr(ib), O.dptr(u),
                            O.dptr(v), O.dptr(uf), O.dptr(vf), C.byref(r))
    assert rel(c.get("u"), u) < 1e-13 and rel(c.get("v"), v) < 1e-13
    assert np.array_equal(c.get("iblank"), np.ones(n * n))
    c.close()


def test_compat_ppe_golden(require_gpu, tmp_path):
    _say("compat PPESolver 52x52")
    xf = O.read_grid_file(os.path.join(GOLDEN, "xgrid.dat2"))
    c = iff.Case(xf, xf, mode=iff.IF_MODE_COMPAT, dt=0.001, re=150.0, ppe_itermax=100000)
    st = c.ppe_solve()
    assert 14000 < st["ppe_sweeps"] < 15500
    iff.write_tecplot(str(tmp_path / "p.dat"), c.xc, c.yc, c.get("p"))
    assert (tmp_path / "p.dat").read_bytes() == open(os.path.join(GOLDEN, "p.dat"), "rb").read()
    c.close()


# ---------------------------------------------------------------------------
# physics mode: FP64 against the oracle
# ---------------------------------------------------------------------------
def _oracle_bodies(bodies):
    out = []
    for b in bodies:
        d = dict(b)
        d["c"], d["s"] = b.get("cth", 1.0), b.get("sth", 0.0)
        out.append(d)
    return out


def _pair(xf, yf, bodies, prec=iff.IF_FP64, **kw):
    prm = dict(dt=kw.get("dt", 0.01), re=kw.get("re", 100.0), conv_factor=0.5, ubc=1.0,
               vbc=0.0, ad_itermax=kw.get("ad_itermax", 25), ad_tol=1e-6,
               ppe_solver=kw.get("ppe_solver", 1), ppe_omega=kw.get("ppe_omega", 1.8),
               ppe_itermax=kw.get("ppe_itermax", 200000), ppe_tol=kw.get("ppe_tol", 1e-9),
               check_every=kw.get("check_every", 10))
    oc = O.PhysCase(xf, yf, _oracle_bodies(bodies), prm)
    gc = iff.Case(xf, yf, precision=prec, bodies=bodies, **prm)
    return oc, gc


def _compare_geometry(oc, gc):
    flags = gc.get("flags").astype(np.uint8)
    assert np.array_equal(flags, oc.flags)
    lst, recs, ng = gc.ghosts()
    assert ng == oc.ng and np.array_equal(lst, oc.glist)
    for g in range(ng):
        a, b = recs[g], oc.gh[g]
        assert list(a.node) == list(b.node) and a.body == b.body
        for k in range(4):
            assert abs(a.wd[k] - b.wd[k]) <= 1e-14 and abs(a.wn[k] - b.wn[k]) <= 1e-14
        assert abs(a.cd - b.cd) <= 1e-14


def _compare_fields(oc, gc, tol):
    for name, ref in (("u", oc.u), ("v", oc.v), ("p", oc.p), ("uf", oc.uf), ("vf", oc.vf)):
        e = rel(gc.get(name), ref)
        assert e <= tol, (name, e)


CASES = {
    "circle-stretched": (stretched_faces(72, 4.0, 1.01), stretched_faces(50, 3.0, 0.99),
                         [iff.circle(1.6, 1.45, 0.45)]),
    "ellipse": (np.linspace(0, 4.0, 97), np.linspace(0, 2.0, 49),
                [iff.ellipse(1.5, 1.0, 0.5, 0.12, np.radians(10.0))]),
    "stars": (np.linspace(0, 4.0, 129), np.linspace(0, 4.0, 129),
              [iff.star(1.3, 1.2, 0.45, [0.05, 0.03, 0.02, 0.06, 0.01], [0.1, 1.2, 2.3, 0.4, 5.0]),
               iff.star(2.8, 2.7, 0.4, [0.02, 0.07, 0.04, 0.01, 0.03], [3.0, 0.5, 1.5, 2.5, 4.0])]),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_physics_fp64_matches_oracle(require_gpu, name):
    xf, yf, bodies = CASES[name]
    _say("physics fp64", name, len(xf) - 1, "x", len(yf) - 1)
    oc, gc = _pair(xf, yf, bodies, dt=0.01)
    _compare_geometry(oc, gc)
    _compare_fields(oc, gc, 1e-14)   # initial state
    for s in range(4):
        o = oc.step(1)[-1]
        g = gc.step(1)
        assert g["ad_iters"] == o[0], (s, g, o)
        assert g["ppe_sweeps"] == o[2], (s, g, o)
        assert abs(g["div_max"] - o[4]) <= 1e-10 * max(o[4], 1e-12) + 1e-18
        _compare_fields(oc, gc, 1e-10)
    gc.close()


@pytest.mark.parametrize("kind,ty,m", [(0, 8, 1), (0, 32, 1), (0, 16, 2), (0, 32, 4), (0, 8, 4),
                                         (1, 16, 1), (1, 32, 2), (1, 64, 2), (1, 128, 4),
                                         (1, 16, 4), (1, 32, 1),
                                         (2, 16, 1), (2, 32, 2), (2, 16, 2), (2, 48, 2), (2, 32, 4),
                                         (2, 48, 4), (3, 24, 2), (3, 48, 4), (3, 32, 8), (3, 48, 8),
                                         (4, 48, 4), (4, 80, 4), (4, 32, 8), (4, 64, 8), (4, 96, 8), (3, 64, 8)])
def test_sweep_variants_are_identical(require_gpu, kind, ty, m):
    """Kernel, tile height and temporal blocking change the launch shape,
    never the result: equal to the oracle's red-black SOR sweep for sweep
    (check_every=13 also exercises partial launches at chunk ends)."""
    xf, yf, bodies = CASES["circle-stretched"]
    _say("sweep variant", kind, ty, m)
    fp32_only = (kind, ty, m) == (4, 96, 8)   # 16 waves x 8 rows: no FP64 instance
    for prec in (iff.IF_FP64, iff.IF_MIXED):
        oc, gc = _pair(xf, yf, bodies, prec=prec, dt=0.01, check_every=13)
        if fp32_only and prec == iff.IF_FP64:
            with pytest.raises(iff.ImmerseFlowError):
                gc.tune(kind, ty, m, 13)
            gc.close()
            continue
        ref = None
        if prec == iff.IF_MIXED:
            ref = iff.Case(xf, yf, precision=prec, bodies=bodies, dt=0.01, re=100.0, ppe_omega=1.8,
                           ppe_itermax=200000, ppe_tol=1e-9, check_every=13)
            ref.tune(0, 16, 1, 13)
        gc.tune(kind, ty, m, 13)
        for _ in range(3):
            o = oc.step(1)[-1]
            g = gc.step(1)
            if prec == iff.IF_FP64:
                assert g["ppe_sweeps"] == o[2]
            else:
                r = ref.step(1)
                assert g["ppe_sweeps"] == r["ppe_sweeps"]
        if prec == iff.IF_FP64:
            _compare_fields(oc, gc, 1e-10)
        else:
            # FP32 sweeps: the kernels agree with each other bit for bit
            for f in ("u", "v", "p", "uf", "vf"):
                assert np.array_equal(gc.get(f), ref.get(f)), f
            ref.close()
        gc.close()


@pytest.mark.parametrize("kind,ty,m", [(0, 32, 4), (1, 16, 1), (1, 64, 2), (1, 128, 4), (2, 32, 2), (2, 48, 4),
                                         (3, 32, 8), (4, 64, 8), (4, 88, 10), (4, 92, 9)])
def test_uniform_grid_form_identical_across_kernels(require_gpu, kind, ty, m):
    """Mixed precision on a grid with one coefficient value per axis: every
    FP32 kernel takes the five-operation update (the register kernel with
    its weight table, the LDS and streaming kernels forming the weights per
    update) and they agree bit for bit with the LDS kernel at its smallest
    shape, sweep counts included."""
    xf, yf, bodies = CASES["stars"]
    _say("uniform-grid form", kind, ty, m)
    kw = dict(precision=iff.IF_MIXED, bodies=bodies, dt=0.01, re=100.0, ppe_omega=1.8, ppe_itermax=200000,
              ppe_tol=1e-9, check_every=13)
    gc, ref = iff.Case(xf, yf, **kw), iff.Case(xf, yf, **kw)
    gc.tune(kind, ty, m, 13)
    ref.tune(0, 16, 1, 13)
    for _ in range(3):
        g, r = gc.step(1), ref.step(1)
        assert g["ppe_sweeps"] == r["ppe_sweeps"] and g["weight_table"] == 1
    for f in ("u", "v", "p", "uf", "vf"):
        assert np.array_equal(gc.get(f), ref.get(f)), f
    gc.close()
    ref.close()


FP32_ONLY_SHAPES = [(4, 96, 8), (4, 92, 9), (4, 88, 10), (4, 84, 11), (4, 80, 12)]


@pytest.mark.parametrize("kind,ty,m", [(2, 16, 2), (2, 32, 4), (2, 48, 1), (3, 40, 2), (3, 64, 4), (3, 32, 8),
                                         (4, 64, 8)] + FP32_ONLY_SHAPES)
def test_register_tile_across_tiles_and_edges(require_gpu, kind, ty, m):
    """Register-tile sweep on a grid several tiles wide: bodies cut by tile
    boundaries and lying next to the domain edge, partial tiles at the east
    and north edges; FP64 equal to the oracle, mixed equal to the LDS kernel
    (the 16-wave tiles of 96 rows and more sweeps per launch are mixed
    precision only)."""
    xf = stretched_faces(300, 6.0, 1.002)
    yf = stretched_faces(140, 2.8, 0.997)
    bodies = [iff.ellipse(2.4, 1.3, 0.45, 0.1, np.radians(20.0)),
              iff.circle(5.72, 0.35, 0.2)]
    _say("register tile edges", kind, ty, m)
    for prec in (iff.IF_MIXED,) if (kind, ty, m) in FP32_ONLY_SHAPES else (iff.IF_FP64, iff.IF_MIXED):
        oc, gc = _pair(xf, yf, bodies, prec=prec, dt=0.005, check_every=9, ppe_tol=1e-8)
        gc.tune(kind, ty, m, 9)
        ref = None
        if prec == iff.IF_MIXED:
            ref = iff.Case(xf, yf, precision=prec, bodies=bodies, dt=0.005, re=100.0, ppe_omega=1.8,
                           ppe_itermax=200000, ppe_tol=1e-8, check_every=9)
            ref.tune(0, 16, 1, 9)
        for _ in range(2):
            if prec == iff.IF_FP64:
                o = oc.step(1)[-1]
                assert gc.step(1)["ppe_sweeps"] == o[2]
            else:
                assert gc.step(1)["ppe_sweeps"] == ref.step(1)["ppe_sweeps"]
        if prec == iff.IF_FP64:
            _compare_fields(oc, gc, 1e-10)
        else:
            for f in ("u", "v", "p", "uf", "vf"):
                assert np.array_equal(gc.get(f), ref.get(f)), f
            ref.close()
        gc.close()


def test_weighted_jacobi_matches_oracle(require_gpu):
    xf, yf, bodies = CASES["circle-stretched"]
    _say("weighted jacobi")
    oc, gc = _pair(xf, yf, bodies, dt=0.01, ppe_solver=0, ppe_omega=0.8, ppe_tol=1e-6,
                   check_every=25)
    for _ in range(2):
        o = oc.step(1)[-1]
        g = gc.step(1)
        assert g["ppe_sweeps"] == o[2]
    _compare_fields(oc, gc, 1e-10)
    gc.close()


# rows of 81 cells: the line solve's partition ends in a block of a single cell
RAGGED = (stretched_faces(81, 4.0, 1.005), np.linspace(0, 2.0, 40), [iff.circle(1.7, 0.9, 0.4)])


@pytest.mark.parametrize("name", ["circle-stretched", "stars", "ragged"])
def test_line_sor_matches_oracle(require_gpu, name):
    """The deck's PPE_Solver 2 (zebra line SOR): FP64 equal to the oracle
    sweep for sweep; mixed within 1e-4 of it."""
    xf, yf, bodies = RAGGED if name == "ragged" else CASES[name]
    _say("line SOR", name)
    oc, gc = _pair(xf, yf, bodies, dt=0.01, ppe_solver=2, ppe_omega=1.5, ppe_tol=1e-9, check_every=7)
    for _ in range(3):
        o = oc.step(1)[-1]
        g = gc.step(1)
        assert g["ppe_sweeps"] == o[2] and g["ad_iters"] == o[0]
        _compare_fields(oc, gc, 1e-10)
    gc.close()
    oc, gm = _pair(xf, yf, bodies, prec=iff.IF_MIXED, dt=0.01, ppe_solver=2, ppe_omega=1.5, ppe_tol=1e-9,
                   check_every=7)
    oc.step(2)
    gm.step(2)
    _compare_fields(oc, gm, 1e-4)
    gm.close()


def test_line_sor_row_width_limit(require_gpu):
    """The line solve stages a row in LDS: rows of more than 64 x 126 = 8064
    interior cells are refused at case creation, and 8064 is accepted."""
    yf = np.linspace(0.0, 1.0, 5)
    with pytest.raises(iff.ImmerseFlowError) as e:
        iff.Case(np.linspace(0.0, 1.0, 8066), yf, dt=1e-3, re=100.0, ppe_solver=iff.IF_PPE_LINE)
    assert e.value.code == iff._abi.IF_ERR_ARG and "8064" in str(e.value)
    c = iff.Case(np.linspace(0.0, 1.0, 8065), yf, dt=1e-3, re=100.0, ppe_solver=iff.IF_PPE_LINE,
                 ppe_tol=1e-8, ppe_itermax=50, check_every=10)
    c.step(1)
    c.close()


def test_batch_mixes_poisson_solvers(require_gpu):
    """Jacobi, red-black SOR and line SOR cases in one batch: each result
    equals the same case alone (the kernels skip the other solvers' cases,
    the checks use each case's own sweeps per launch)."""
    xf, yf, bodies = CASES["circle-stretched"]
    _say("batch of three Poisson solvers")
    specs = [(0, 0.8), (1, 1.8), (2, 1.5)]
    kw = dict(dt=0.01, re=100.0, bodies=bodies, ppe_tol=1e-8, check_every=12)
    for prec in (iff.IF_FP64, iff.IF_MIXED):
        mix = iff.Batch(prec)
        slots = [mix.add(xf, yf, 2, ppe_solver=sv, ppe_omega=om, **kw) for sv, om in specs]
        mix.step(2)
        for k, (sv, om) in enumerate(specs):
            alone = iff.Batch(prec)
            sl = alone.add(xf, yf, 2, ppe_solver=sv, ppe_omega=om, **kw)
            alone.step(2)
            assert alone.state(sl)[2]["ppe_sweeps"] == mix.state(slots[k])[2]["ppe_sweeps"]
            for f in ("u", "v", "p"):
                assert np.array_equal(alone.get(sl, f), mix.get(slots[k], f)), (prec, sv, f)
            alone.close()
        mix.close()


@pytest.mark.parametrize("prec", [iff.IF_FP64, iff.IF_MIXED])
def test_graph_replay_equals_direct_launches(require_gpu, prec):
    """Poisson chunks replayed from a captured hipGraph (the default) give
    the same bits as the kernels launched one by one: over several steps
    (many replays, eager kernels and host synchronisation between them),
    after a launch-shape change (recapture) and after a case joins the
    batch (new case list and tile tables, recapture)."""
    xf, yf, bodies = CASES["circle-stretched"]
    xf2, yf2, bodies2 = CASES["ellipse"]
    _say("graph replay vs direct launches", prec)
    kw = dict(dt=0.01, re=100.0, ppe_omega=1.8, ppe_tol=1e-8, check_every=16)
    runs = []
    for graphs in (True, False):
        b = iff.Batch(prec)
        b.graphs(graphs)
        s1 = b.add(xf, yf, 6, bodies=bodies, **kw)
        b.step(2)
        b.tune(3, 24, 2, 12)
        b.step(1)
        s2 = b.add(xf2, yf2, 3, bodies=bodies2, **kw)
        b.step(3)
        runs.append({(s, f): b.get(s, f) for s in (s1, s2) for f in ("u", "v", "p", "uf", "vf")})
        runs[-1]["sweeps"] = [b.state(s)[2]["total_ppe_sweeps"] for s in (s1, s2)]
        b.close()
    assert runs[0]["sweeps"] == runs[1]["sweeps"]
    for k in runs[0]:
        if k != "sweeps":
            assert np.array_equal(runs[0][k], runs[1][k]), k


@pytest.mark.parametrize("prec,kind,ty,m,K,graphs", [
    (iff.IF_MIXED, 4, 96, 8, 40, True), (iff.IF_MIXED, 4, 64, 8, 13, True), (iff.IF_MIXED, 3, 32, 8, 40, False),
    (iff.IF_MIXED, 4, 48, 4, 12, True), (iff.IF_MIXED, 3, 32, 4, 10, True),
    (iff.IF_FP64, 4, 80, 4, 12, True), (iff.IF_FP64, 4, 48, 4, 13, False)])
def test_persistent_chunks_equal_launch_per_stage(require_gpu, prec, kind, ty, m, K, graphs):
    """A chunk run as one persistent launch (tiles kept in registers, halo
    strips exchanged between neighbouring tiles between stages) gives the
    same bits as one launch per stage, in FP32 error sweeps and in FP64
    sweeps: in a batch with a point-Jacobi case (its blocks only advance
    their stage counters), a case whose sweep budget ends inside a chunk, a
    heaving body (flags, ghosts and weights rebuilt every step), and chunks
    that are not a whole number of stages."""
    xf, yf, bodies = CASES["circle-stretched"]
    xf2, yf2, bodies2 = CASES["ellipse"]
    _say("persistent vs per-stage launches", prec, kind, ty, m, K, graphs)
    kw = dict(dt=0.01, re=100.0, ppe_omega=1.8, ppe_tol=1e-8, check_every=K)
    runs = []
    for pers in (True, False):
        b = iff.Batch(prec)
        b.tune(kind, ty, m, K)
        b.graphs(graphs)
        b.persistent(pers)
        sl = [b.add(xf, yf, 3, bodies=bodies, **kw),
              b.add(xf2, yf2, 3, bodies=bodies2, **kw),
              b.add(xf2, yf2, 3, bodies=bodies2, **dict(kw, ppe_itermax=3 * K + 5)),
              b.add(xf, yf, 3, bodies=bodies, **dict(kw, ppe_solver=0, ppe_omega=0.8)),
              b.add(xf2, yf2, 3, bodies=[iff.heaving_circle(2.0, 1.0, 0.3, 0.2, 0.5)], **kw)]
        b.step(3)
        runs.append({(k, f): b.get(s, f) for k, s in enumerate(sl) for f in ("u", "v", "p", "uf", "vf")})
        runs[-1]["sweeps"] = [b.state(s)[2]["total_ppe_sweeps"] for s in sl]
        b.close()
    assert runs[0]["sweeps"] == runs[1]["sweeps"]
    for k in runs[0]:
        if k != "sweeps":
            assert np.array_equal(runs[0][k], runs[1][k]), k


@pytest.mark.parametrize("prec", [iff.IF_MIXED, iff.IF_MIXED_V32])
def test_mixed_precision_against_fp64_oracle(require_gpu, prec):
    xf, yf, bodies = CASES["circle-stretched"]
    _say("mixed precision", prec)
    oc, gc = _pair(xf, yf, bodies, prec=prec, dt=0.01, ppe_tol=1e-9, check_every=10)
    for _ in range(4):
        oc.step(1)
        g = gc.step(1)
        assert g["div_max"] <= 2e-9
    for name, ref in (("u", oc.u), ("v", oc.v), ("uf", oc.uf), ("vf", oc.vf)):
        assert rel(gc.get(name), ref) <= 1e-4, name
    # pressure up to its additive constant (volume-weighted mean over fluid)
    fl = (oc.flags & 1).astype(bool)
    W = np.outer(oc.dy, oc.dx).ravel()
    pg, po = gc.get("p"), oc.p
    inner = np.zeros_like(fl)
    F2 = fl.reshape(oc.ny, oc.nx)
    F2[[0, -1], :] = False
    F2[:, [0, -1]] = False
    m = F2.ravel()
    pg = pg - np.sum(W[m] * pg[m]) / np.sum(W[m])
    po = po - np.sum(W[m] * po[m]) / np.sum(W[m])
    assert rel(pg[m], po[m]) <= 1e-4
    gc.close()


# ---------------------------------------------------------------------------
# batching: a case's result does not depend on what else is resident
# ---------------------------------------------------------------------------
def test_batch_alone_equals_in_mix(require_gpu):
    specs = [
        (np.linspace(0, 3, 61), np.linspace(0, 2, 41), [iff.circle(1.0, 1.0, 0.3)], 0.01, 120.0),
        (stretched_faces(90, 4, 1.01), stretched_faces(70, 3, 1.0), [iff.ellipse(2, 1.5, 0.4, 0.1, 0.3)], 0.008, 300.0),
        (np.linspace(0, 2, 33), np.linspace(0, 2, 33), [iff.circle(0.8, 1.0, 0.25)], 0.02, 50.0),
    ]
    steps = [3, 5, 4]
    _say("batch alone vs mix")
    kw = dict(ppe_tol=1e-8, check_every=10, ppe_omega=1.85)
    mix = iff.Batch(iff.IF_FP64)
    mix.tune(0, 16, 2, 10)
    slots = [mix.add(xf, yf, n, dt=dt, re=re, bodies=b, **kw)
             for (xf, yf, b, dt, re), n in zip(specs, steps)]
    while any(mix.state(s)[0] == 0 for s in slots):
        mix.step(1)
    for k, ((xf, yf, b, dt, re), n) in enumerate(zip(specs, steps)):
        st, done, _ = mix.state(slots[k])
        assert st == 1 and done == n
        alone = iff.Batch(iff.IF_FP64)
        alone.tune(0, 16, 2, 10)
        s = alone.add(xf, yf, n, dt=dt, re=re, bodies=b, **kw)
        alone.step(n)
        for f in ("u", "v", "p", "uf", "vf"):
            assert np.array_equal(alone.get(s, f), mix.get(slots[k], f)), (k, f)
        alone.close()
    mix.close()


@pytest.mark.parametrize("prec", [iff.IF_FP64, iff.IF_MIXED])
def test_autotune_leaves_state_and_matches_manual_tune(require_gpu, prec):
    """The autotuner times launch shapes on the live fields but leaves the
    solution and loop state as they were: afterwards the case continues
    exactly like one given the chosen shape and interval by hand."""
    xf, yf, bodies = CASES["ellipse"]
    _say("autotune", prec)
    kw = dict(precision=prec, bodies=bodies, dt=0.01, re=100.0, ppe_omega=1.8, ppe_itermax=200000,
              ppe_tol=1e-9, check_every=10)
    a, b = iff.Case(xf, yf, **kw), iff.Case(xf, yf, **kw)
    ncand = 13 if prec == iff.IF_FP64 else 16   # the 16-wave tiles of 96/92/88 rows are FP32 only
    r0 = a.autotune()                       # before any step: interval unchanged
    assert r0["typical_sweeps"] == 0 and r0["check_every"] == 10 and r0["candidates"] == ncand
    a.step(2)
    b.step(2)
    r = a.autotune()
    assert r["candidates"] == ncand and r["typical_sweeps"] > 0
    assert r["check_every"] % r["sweeps_per_launch"] == 0 and 4 <= r["check_every"] <= 1024
    assert r["sweep_ns"] > 0 and r["check_sweeps"] > 0
    for f in ("u", "v", "p", "uf", "vf"):
        assert np.array_equal(a.get(f), b.get(f)), f
    b.tune(r["kernel"], r["tile_y"], r["sweeps_per_launch"], r["check_every"])
    for _ in range(2):
        sa, sb = a.step(1), b.step(1)
        assert sa["ppe_sweeps"] == sb["ppe_sweeps"] and sa["ad_iters"] == sb["ad_iters"]
    for f in ("u", "v", "p", "uf", "vf"):
        assert np.array_equal(a.get(f), b.get(f)), f
    a.close()
    b.close()


def test_batch_autotune(require_gpu):
    specs = [(np.linspace(0, 3, 61), np.linspace(0, 2, 41), [iff.circle(1.0, 1.0, 0.3)], 0.01, 120.0),
             (stretched_faces(150, 4, 1.005), stretched_faces(90, 3, 1.0), [iff.ellipse(2, 1.5, 0.4, 0.1, 0.3)],
              0.008, 300.0)]
    _say("batch autotune")
    kw = dict(ppe_tol=1e-8, check_every=10, ppe_omega=1.85)
    bs = []
    for _ in range(2):
        bt = iff.Batch(iff.IF_MIXED)
        sl = [bt.add(xf, yf, 6, dt=dt, re=re, bodies=bo, **kw) for xf, yf, bo, dt, re in specs]
        bt.step(1)
        bs.append((bt, sl))
    r = bs[0][0].autotune()
    assert r["candidates"] == 16 and r["typical_sweeps"] > 0
    bs[1][0].tune(r["kernel"], r["tile_y"], r["sweeps_per_launch"], r["check_every"])
    for bt, _ in bs:
        bt.step(2)
    for k in range(len(specs)):
        for f in ("u", "v", "p"):
            assert np.array_equal(bs[0][0].get(bs[0][1][k], f), bs[1][0].get(bs[1][1][k], f)), (k, f)
    for bt, _ in bs:
        bt.close()


def test_batch_budget_admission(require_gpu):
    xf = np.linspace(0, 1, 65)
    need = iff.Batch.footprint(xf, xf, dt=0.01, re=100.0)
    b = iff.Batch(iff.IF_FP64, hbm_budget=int(need * 1.5))
    assert b.add(xf, xf, 1, dt=0.01, re=100.0) is not None
    assert b.add(xf, xf, 1, dt=0.01, re=100.0) is None   # refused: over budget
    b.close()


# ---------------------------------------------------------------------------
# moving bodies: reclassification, fresh cells and ghost rebuild per step
# ---------------------------------------------------------------------------
def test_heaving_cylinder_matches_oracle(require_gpu):
    import math
    xf = np.linspace(0, 4, 81)
    y0, A, f = 2.0, 0.5, 0.2
    w = 2 * math.pi * f
    dt = 0.02
    _say("heaving cylinder")
    prm = dict(dt=dt, re=100.0, conv_factor=0.5, ubc=1.0, vbc=0.0, ad_itermax=25, ad_tol=1e-6,
               ppe_solver=1, ppe_omega=1.8, ppe_itermax=100000, ppe_tol=1e-8, check_every=10)
    # the body's state at t0 = 0
    oc = O.PhysCase(xf, xf, [dict(kind=O.CIRCLE, xc=1.5, yc=y0, r=0.4, vb=A * w)], prm)
    gc = iff.Case(xf, xf, bodies=[iff.heaving_circle(1.5, y0, 0.4, A, f)], **prm)
    _compare_geometry(oc, gc)
    _compare_fields(oc, gc, 1e-14)
    for n in range(6):
        oc.heave({0: (y0, A, f)}, (n + 1) * dt)
        o = oc.step(1)[-1]
        g = gc.step(1)
        _compare_geometry(oc, gc)
        assert g["ppe_sweeps"] == o[2] and g["ad_iters"] == o[0], (n, g, o)
        _compare_fields(oc, gc, 1e-10)
    gc.close()


@pytest.mark.parametrize("kind,ty,m", [(3, 32, 4), (2, 32, 2), (4, 48, 4), (4, 80, 4), (3, 64, 8)])
def test_fp64_weight_table_matches_oracle(require_gpu, kind, ty, m):
    """FP64 on a grid with one value per Poisson coefficient (spacing 1/32):
    the register-tile kernel looks each cell's weight up in a table formed
    with the oracle's expression instead of dividing on every launch, per
    stage and as persistent chunks; sweep counts equal to the oracle's and
    fields to 1e-10, with a star (cells of every face pattern) and with a
    heaving body (patterns change every step).  A stretched grid takes the
    division."""
    import math
    xf, yf, bodies = CASES["stars"]
    _say("fp64 weight table", kind, ty, m)
    oc, gc = _pair(xf, yf, bodies, dt=0.01, check_every=12)
    gc.tune(kind, ty, m, 12)
    for _ in range(2):
        o = oc.step(1)[-1]
        g = gc.step(1)
        assert g["weight_table"] == 1 and g["ppe_sweeps"] == o[2], (g, o)
    _compare_fields(oc, gc, 1e-10)
    gc.close()
    y0, A, f, dt = 2.0, 0.5, 0.2, 0.02
    prm = dict(dt=dt, re=100.0, conv_factor=0.5, ubc=1.0, vbc=0.0, ad_itermax=25, ad_tol=1e-6,
               ppe_solver=1, ppe_omega=1.8, ppe_itermax=100000, ppe_tol=1e-8, check_every=12)
    oc = O.PhysCase(xf, xf, [dict(kind=O.CIRCLE, xc=1.5, yc=y0, r=0.4, vb=A * 2 * math.pi * f)], prm)
    gc = iff.Case(xf, xf, bodies=[iff.heaving_circle(1.5, y0, 0.4, A, f)], **prm)
    gc.tune(kind, ty, m, 12)
    for n in range(3):
        oc.heave({0: (y0, A, f)}, (n + 1) * dt)
        o = oc.step(1)[-1]
        g = gc.step(1)
        assert g["weight_table"] == 1 and g["ppe_sweeps"] == o[2], (n, g, o)
        _compare_fields(oc, gc, 1e-10)
    gc.close()
    xs, ys, bs = CASES["circle-stretched"]
    gs = iff.Case(xs, ys, bodies=bs, dt=0.01, re=100.0, ppe_omega=1.8, ppe_tol=1e-9, check_every=12)
    gs.tune(kind, ty, m, 12)
    assert gs.step(1)["weight_table"] == 0
    gs.close()


def test_mixed_weights_follow_moving_flags(require_gpu):
    """Mixed precision with a heaving body: the register-tile kernel loads
    the FP32 SOR weights formed once per step from the flags, the LDS-tiled
    kernel forms them from the flags itself.  Both must give the same bits
    step after step while the body moves and cells change status (stale
    weights would also pass an accuracy test: the FP64 defect correction
    hides them)."""
    xf = np.linspace(0, 4, 81)
    kw = dict(precision=iff.IF_MIXED, bodies=[iff.heaving_circle(1.5, 2.0, 0.4, 0.5, 0.4)], dt=0.02, re=100.0,
              ppe_omega=1.8, ppe_itermax=100000, ppe_tol=1e-8, check_every=16)
    _say("mixed weights with a heaving body")
    a, b = iff.Case(xf, xf, **kw), iff.Case(xf, xf, **kw)
    a.tune(4, 64, 8, 16)
    b.tune(0, 16, 2, 16)
    moved = 0
    f0 = a.get("flags")
    for n in range(6):
        sa, sb = a.step(1), b.step(1)
        assert sa["ppe_sweeps"] == sb["ppe_sweeps"], (n, sa, sb)
        for f in ("u", "v", "p", "uf", "vf", "flags"):
            assert np.array_equal(a.get(f), b.get(f)), (n, f)
        moved += int(np.count_nonzero(a.get("flags") != f0))
        f0 = a.get("flags")
    assert moved > 0  # cells changed status along the way
    a.close()
    b.close()


def test_mixed_weight_table_equals_per_cell_weights(require_gpu):
    """On a grid whose FP32 Poisson coefficients are each one value (spacing
    1/16) the register-tile kernel looks each cell's weight up by its face
    flags.  Run alone, a static and a heaving case take that form; in a
    batch with a stretched case every case loads per-cell weights.  Both
    forms, and the LDS-tiled kernel (which forms the weights itself), must
    give the same bits step after step."""
    xu = np.linspace(0.0, 4.0, 65)
    xs = stretched_faces(70, 4.0, 1.01)
    kw = dict(dt=0.02, re=100.0, ppe_omega=1.8, ppe_itermax=100000, ppe_tol=1e-8, check_every=16)
    bodies = {"static": [iff.circle(1.5, 2.0, 0.4)], "heave": [iff.heaving_circle(1.5, 2.0, 0.4, 0.5, 0.4)]}
    _say("mixed weight table vs per-cell weights")
    for name, bd in bodies.items():
        alone = iff.Case(xu, xu, precision=iff.IF_MIXED, bodies=bd, **kw)
        lds = iff.Case(xu, xu, precision=iff.IF_MIXED, bodies=bd, **kw)
        alone.tune(4, 64, 8, 16)
        lds.tune(0, 16, 2, 16)
        b = iff.Batch(iff.IF_MIXED)
        b.tune(4, 64, 8, 16)
        s0 = b.add(xu, xu, 4, bodies=bd, **kw)
        b.add(xs, xs, 4, bodies=[iff.circle(2.0, 2.0, 0.5)], **kw)
        for n in range(4):
            sa, sl = alone.step(1), lds.step(1)
            assert sa["weight_table"] == 1 and sl["weight_table"] == 1, (name, n)
            assert sa["ppe_sweeps"] == sl["ppe_sweeps"], (name, n)
        b.step(4)
        st = b.state(s0)[2]
        assert st["weight_table"] == 0 and st["total_ppe_sweeps"] > 0, name
        for f in ("u", "v", "p", "uf", "vf", "flags"):
            assert np.array_equal(alone.get(f), lds.get(f)), (name, f)
            assert np.array_equal(alone.get(f), b.get(s0, f)), (name, f)
        alone.close()
        lds.close()
        b.close()


def test_batch_moving_and_static_alone_equals_mix(require_gpu):
    xs = np.linspace(0, 4, 65)
    kw = dict(dt=0.02, re=150.0, ppe_tol=1e-8, check_every=10, ppe_omega=1.8)
    cases = [([iff.heaving_circle(1.5, 2.0, 0.4, 0.5, 0.2)], 5),
             ([iff.circle(2.0, 2.0, 0.5)], 4),
             ([iff.heaving_circle(2.5, 1.8, 0.3, 0.3, 0.5)], 6)]
    _say("batch moving + static")
    mix = iff.Batch(iff.IF_FP64)
    slots = [mix.add(xs, xs, n, bodies=b, **kw) for b, n in cases]
    while any(mix.state(s)[0] == 0 for s in slots):
        mix.step(1)
    for k, (b, n) in enumerate(cases):
        alone = iff.Batch(iff.IF_FP64)
        s = alone.add(xs, xs, n, bodies=b, **kw)
        alone.step(n)
        for fld in ("u", "v", "p", "uf", "vf", "flags"):
            assert np.array_equal(alone.get(s, fld), mix.get(slots[k], fld)), (k, fld)
        alone.close()
    mix.close()
