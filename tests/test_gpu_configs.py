"""GPU parity at the workloads' full sizes.

* 256^2 cylinder (FP64, Poisson stop 1e-8): equal to the oracle step for
  step (sweep counts identical, fields to 1e-10).
* 2048x1024 stretched ellipse: one FP64 step equal to the oracle's; the
  mixed path within 1e-5 of the FP64 path at the same stop level (1e-4
  with FP32 velocity stencils).
* 4096^2 with eight star bodies: flags and ghost list bit-equal to the
  oracle's geometry; one mixed step within 1e-5 of the FP64 path, both at
  max|div| below the stop level.
* A sample of the 256-case batch (sizes 128^2..1024^2): each case's result
  in the batch equals the case run alone.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from test_gpu_parity import _compare_fields, _compare_geometry, _oracle_bodies, rel

import immerseflow_amd as iff
from immerseflow_amd import configs

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-test]", *a, flush=True)


def _prm(cfg, omega, ppe_tol, check_every):
    return dict(dt=cfg["dt"], re=cfg["re"], conv_factor=0.5, ubc=1.0, vbc=0.0, ad_itermax=25, ad_tol=1e-6,
                ppe_solver=1, ppe_omega=omega, ppe_itermax=200000, ppe_tol=ppe_tol, check_every=check_every)


def _default_omega(cfg):
    n = max(len(cfg["xf"]), len(cfg["yf"])) - 1
    return 2.0 / (1.0 + math.sin(math.pi / n))


def test_c1_cylinder_256_fp64_matches_oracle(require_gpu):
    cfg = configs.c1()
    prm = _prm(cfg, _default_omega(cfg), cfg["ppe_tol"], 20)
    _say("C1 256^2 fp64 vs oracle")
    oc = O.PhysCase(cfg["xf"], cfg["yf"], _oracle_bodies(cfg["bodies"]), prm)
    gc = iff.Case(cfg["xf"], cfg["yf"], precision=iff.IF_FP64, bodies=cfg["bodies"], **prm)
    _compare_geometry(oc, gc)
    for s in range(3):
        o = oc.step(1)[-1]
        g = gc.step(1)
        assert g["ad_iters"] == o[0] and g["ppe_sweeps"] == o[2], (s, g, o)
        assert g["div_max"] <= cfg["ppe_tol"]
        _compare_fields(oc, gc, 1e-10)
    gc.close()


def test_c2_ellipse_2048x1024_fp64_step_matches_oracle(require_gpu):
    cfg = configs.c2()
    prm = _prm(cfg, 1.998, cfg["ppe_tol"], 160)
    _say("C2 2048x1024 fp64 vs oracle, 1 step")
    oc = O.PhysCase(cfg["xf"], cfg["yf"], _oracle_bodies(cfg["bodies"]), prm)
    gc = iff.Case(cfg["xf"], cfg["yf"], precision=iff.IF_FP64, bodies=cfg["bodies"], **prm)
    _compare_geometry(oc, gc)
    o = oc.step(1)[-1]
    g = gc.step(1)
    assert g["ad_iters"] == o[0] and g["ppe_sweeps"] == o[2], (g, o)
    _compare_fields(oc, gc, 1e-10)
    gc.close()


def _mixed_vs_fp64(cfg, prm, steps, bounds):
    """bounds: {precision: tolerance on u, v, p against the FP64 path}"""
    out = {}
    for prec in [iff.IF_FP64] + list(bounds):
        c = iff.Case(cfg["xf"], cfg["yf"], precision=prec, bodies=cfg["bodies"], **prm)
        st = c.step(steps)
        assert st["max_div"] <= prm["ppe_tol"], (prec, st)
        out[prec] = {f: c.get(f) for f in ("u", "v", "p")}
        c.close()
    for prec, tol in bounds.items():
        for f in ("u", "v", "p"):
            e = rel(out[prec][f], out[iff.IF_FP64][f])
            _say("  precision", prec, "vs fp64", f, e)
            assert e <= tol, (prec, f, e)


def test_c2_mixed_vs_fp64(require_gpu):
    """IF_MIXED (FP32 Poisson error sweeps) within 1e-5 of the FP64 path;
    IF_MIXED_V32 (FP32 velocity stencils too) within 1e-4, the bound the
    pressure's sensitivity to FP32 velocity rounding allows."""
    cfg = configs.c2()
    _say("C2 mixed vs fp64, 2 steps")
    _mixed_vs_fp64(cfg, _prm(cfg, 1.998, cfg["ppe_tol"], 160), 2, {iff.IF_MIXED: 1e-5, iff.IF_MIXED_V32: 1e-4})


@pytest.mark.parametrize("prec,shape", [(iff.IF_MIXED, (4, 96, 8, 240)), (iff.IF_FP64, (4, 80, 4, 160))])
def test_c2_persistent_chunks_equal_launch_per_stage(require_gpu, prec, shape):
    """The 2048x1024 case at the mixed default shape (242 tiles) and at the
    FP64 persistent shape (247 tiles), one block per CU, halo strips
    crossing every XCD boundary: persistent chunks give the same bits and
    sweep counts as one launch per stage."""
    cfg = configs.c2()
    prm = _prm(cfg, 1.998, cfg["ppe_tol"], shape[3])
    _say("C2 persistent vs per-stage, 2 steps", prec, shape)
    out = []
    for pers in (True, False):
        c = iff.Case(cfg["xf"], cfg["yf"], precision=prec, bodies=cfg["bodies"], **prm)
        c.tune(*shape)
        c.persistent(pers)
        st = c.step(2)
        out.append((st["total_ppe_sweeps"], {f: c.get(f) for f in ("u", "v", "p", "uf", "vf")}))
        c.close()
    assert out[0][0] == out[1][0]
    for f in out[0][1]:
        assert np.array_equal(out[0][1][f], out[1][1][f]), f


def test_c3_stars_4096_geometry_and_mixed(require_gpu):
    cfg = configs.c3()
    prm = _prm(cfg, 0.0, cfg["ppe_tol"], 160)
    _say("C3 4096^2 geometry vs oracle")
    fl, bid = O.classify(*[O.metric(f)[0] for f in (cfg["xf"], cfg["yf"])], O.BodySet(_oracle_bodies(cfg["bodies"])))
    gc = iff.Case(cfg["xf"], cfg["yf"], precision=iff.IF_MIXED, bodies=cfg["bodies"], **prm)
    assert np.array_equal(gc.get("flags").astype(np.uint8), fl)
    xc, _ = O.metric(cfg["xf"])
    yc, _ = O.metric(cfg["yf"])
    lst, _, ng = O.ghost_records(xc, yc, fl, bid, O.BodySet(_oracle_bodies(cfg["bodies"])))
    glst, _, gng = gc.ghosts()
    assert gng == ng and np.array_equal(glst, lst)
    gc.close()
    _say("C3 mixed vs fp64, 1 step")
    _mixed_vs_fp64(cfg, prm, 1, {iff.IF_MIXED: 1e-5})


def test_c4_sample_batch_equals_alone(require_gpu):
    cases = configs.c4_cases(256)[::43]   # 6 cases, 128^2..1024^2
    _say("C4 sample batch vs alone", [len(c["xf"]) - 1 for c in cases])
    kw = dict(ppe_tol=1e-6, check_every=160)
    mix = iff.Batch(iff.IF_MIXED)
    slots = [mix.add(c["xf"], c["yf"], 2, dt=c["dt"], re=c["re"], bodies=c["bodies"], **kw) for c in cases]
    mix.step(2)
    for k, c in enumerate(cases):
        assert mix.state(slots[k])[0] == 1
        alone = iff.Batch(iff.IF_MIXED)
        s = alone.add(c["xf"], c["yf"], 2, dt=c["dt"], re=c["re"], bodies=c["bodies"], **kw)
        alone.step(2)
        assert alone.state(s)[2]["max_div"] <= 1e-6
        for f in ("u", "v", "p"):
            assert np.array_equal(alone.get(s, f), mix.get(slots[k], f)), (k, f)
        alone.close()
    mix.close()
