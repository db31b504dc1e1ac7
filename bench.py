#!/usr/bin/env python3
"""Throughput of the fractional step on one MI355X.

Workload: the single large case -- an ellipse (chord 1, thickness ratio 0.1)
at 10 deg incidence, Re = 300, on a 2048 x 1024 grid (uniform h = 0.005 in a
4 x 2 box around the body, stretched to a 30 x 20 domain), dt = 0.5 h.  A
step is one full fractional step (predictor, divergence RHS, Poisson to
max|div| <= 1e-6, projection, divergence check) on device-resident fields.

value = interior cells x steps / wall seconds (Mcell-steps/s) of the
mixed-precision path (FP32 Poisson error sweeps, FP64 residual and defect
correction, FP64 velocity passes, faces and divergence); the FP64 path and
the FP32-velocity variant (IF_MIXED_V32) are timed beside it, each with its
error against the FP64 path.  roofline: the Poisson sweep kernel (the dominant kernel), algorithmic
bytes per launch / mean launch time from HIP events recorded by the library
on its stream during the timed steps.  cpu_baseline: the C oracle (the same
per-cell routines, OpenMP) timed for one step from the GPU's state.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--kernel", type=int, default=4,
                    help="0 LDS-tiled, 1 register-streaming, 2/3/4 register-tile (4/8/16 waves)")
    ap.add_argument("--tile-y", type=int, default=96)
    ap.add_argument("--sweeps-per-launch", type=int, default=8)
    ap.add_argument("--check-every", type=int, default=480)
    ap.add_argument("--fp64-kernel", type=int, default=4)
    ap.add_argument("--fp64-tile-y", type=int, default=80)
    ap.add_argument("--fp64-sweeps-per-launch", type=int, default=4)
    ap.add_argument("--fp64-check-every", type=int, default=160)
    ap.add_argument("--omega", type=float, default=1.998)
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch the Poisson sweep kernels one by one instead of replaying a hipGraph")
    ap.add_argument("--no-persistent", action="store_true",
                    help="one Poisson launch per sweeps-per-launch stage instead of one per check interval")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=2)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-c3", action="store_true")
    ap.add_argument("--no-v32", action="store_true")
    ap.add_argument("--no-line", action="store_true")
    ap.add_argument("--line-cases", default="c1,c2", help="line-SOR workloads to time (c1, c2)")
    ap.add_argument("--c3-steps", type=int, default=2)
    ap.add_argument("--serve-requests", type=int, default=32)
    return ap.parse_args()


def kernel_name(kind, ty, m, t, pers=False):
    """The Poisson sweep kernel a tune() choice launches (as rocprofv3 names it)."""
    if kind == 0:
        return "k_rbsor<%s, %d, %d>" % (t, ty, m)
    if kind == 1:
        return "k_rbsor_stream<%s, %d, %d>" % (t, m, ty)
    nw = {2: 4, 3: 8, 4: 16}[kind]
    return "k_rbsor_reg<%s, %d, %d, %d, %s>" % (t, m, nw, (ty + 4 * m) // nw, "true" if pers else "false")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def rel_err(a, b):
    """max |a - b| / max(1, max |b|): the error of a mixed-precision field
    against the FP64 field of the same case and step."""
    import numpy as np
    return float(np.max(np.abs(a - b))) / max(1.0, float(np.max(np.abs(b))))


def fields_err(get_mixed, get_fp64):
    """u, v, p errors; p also after removing the difference of the two
    fields' means (the pressure of the all-Neumann problem is fixed only up
    to a constant, which neither solve pins)."""
    out = {f: rel_err(get_mixed(f), get_fp64(f)) for f in ("u", "v", "p")}
    pm, pf = get_mixed("p"), get_fp64("p")
    out["p_mod_const"] = rel_err(pm - pm.mean(), pf - pf.mean())
    return out


def make_case(iff, cfg, precision, a):
    if precision == iff.IF_FP64:
        shape = (a.fp64_kernel, a.fp64_tile_y, a.fp64_sweeps_per_launch, a.fp64_check_every)
    else:
        shape = (a.kernel, a.tile_y, a.sweeps_per_launch, a.check_every)
    c = iff.Case(cfg["xf"], cfg["yf"], precision=precision, dt=cfg["dt"], re=cfg["re"],
                 bodies=cfg["bodies"], ad_itermax=25, ad_tol=1e-6, ppe_solver=iff.IF_PPE_RBSOR,
                 ppe_omega=a.omega, ppe_itermax=200000, ppe_tol=cfg["ppe_tol"],
                 check_every=shape[3])
    c.tune(*shape)
    c.graphs(not a.no_graphs)
    c.persistent(not a.no_persistent)
    return c


def timed_run(case, warmup, steps, cells, label):
    for k in range(warmup):
        st = case.step(1)
        log(label, "warmup", k, "sweeps", st["ppe_sweeps"], "div", "%.2e" % st["div_max"])
    case.sweep_timing(reset=True)
    t0 = time.perf_counter()
    st = case.step(steps)
    t1 = time.perf_counter()
    ms, launches = case.sweep_timing(reset=True)
    dt = t1 - t0
    log(label, "timed", steps, "steps in %.3f s" % dt, "last sweeps", st["ppe_sweeps"],
        "total sweeps", st["total_ppe_sweeps"], "max div %.2e" % st["max_div"])
    return dict(seconds=dt, stats=st, sweep_ms=ms, sweep_launches=launches,
                mcs=cells * steps / dt / 1e6)


def cpu_baseline(cfg, gpu_case, st, omega):
    """One oracle step (FP64, OpenMP) from the GPU case's state."""
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    prm = dict(dt=cfg["dt"], re=cfg["re"], conv_factor=0.5, ubc=1.0, vbc=0.0, ad_itermax=25,
               ad_tol=1e-6, ppe_solver=1, ppe_omega=0.0, ppe_itermax=max(1, 2 * st["ppe_sweeps"]),
               ppe_tol=cfg["ppe_tol"], check_every=20)
    bodies = []
    for b in cfg["bodies"]:
        d = dict(b)
        d["c"], d["s"] = b.get("cth", 1.0), b.get("sth", 0.0)
        bodies.append(d)
    oc = O.PhysCase(cfg["xf"], cfg["yf"], bodies, prm)
    oc.prm.ppe_omega = omega
    for name in ("u", "v", "p", "uf", "vf"):
        getattr(oc, name)[:] = gpu_case.get(name)
    t0 = time.perf_counter()
    out = oc.step(1)[-1]
    t1 = time.perf_counter()
    cells = (oc.nx - 2) * (oc.ny - 2)
    return dict(value=round(cells / (t1 - t0) / 1e6, 4), unit="Mcell-steps/s", cores=threads,
                kind="port",
                sample="1 FP64 step of the same case from the GPU state (%d AD sweeps, %d "
                       "Poisson sweeps, %.1f s)" % (out[0], out[2], t1 - t0))


def batched_c4(iff, configs, a):
    """The batched workload (256 seeded uneven cases, 128^2..1024^2) in one
    resident batch, mixed precision, against the plain FP64 single-case path
    (library defaults, one case at a time) on a sample of the same cases.
    Step 1 runs untimed on both; the batch is then autotuned (launch shape
    and check interval, timed separately) and steps 2 .. 1+--batch-steps are
    timed on both, expressed per 200-step case."""
    cases = configs.c4_cases(256)
    b = iff.Batch(iff.IF_MIXED, hbm_budget=0)
    b.tune(a.kernel, a.tile_y, a.sweeps_per_launch, a.check_every)
    b.graphs(not a.no_graphs)
    b.persistent(not a.no_persistent)
    t0 = time.perf_counter()
    slots = [b.add(c["xf"], c["yf"], c["steps"], dt=c["dt"], re=c["re"], bodies=c["bodies"],
                   ppe_tol=1e-6, check_every=a.check_every) for c in cases]
    t_setup = time.perf_counter() - t0
    b.step(1)
    t0 = time.perf_counter()
    tr = b.autotune()
    t_tune = time.perf_counter() - t0
    log("batch autotune:", tr, "%.3f s" % t_tune)
    i0 = b.info()
    t0 = time.perf_counter()
    b.step(a.batch_steps)
    t_b = time.perf_counter() - t0
    info = b.info()
    divs = [b.state(s)[2]["max_div"] for s in slots]
    cells = sum((len(c["xf"]) - 1) * (len(c["yf"]) - 1) for c in cases)
    per_step = t_b / a.batch_steps
    cps = len(cases) / (200.0 * per_step)
    ms = info["sweep_ms"] - i0["sweep_ms"]
    gbs = (info["cell_sweeps"] - i0["cell_sweeps"]) * 13.0 / (ms * 1e-3) / 1e9 if ms > 0 else None
    # plain FP64 single-case path on every 32nd case, run one after another
    # (same steps as the batch: the sampled cases' mixed results are compared
    # with these FP64 results field by field)
    tot = 0.0
    sample = list(range(0, len(cases), 32))
    err = {"u": 0.0, "v": 0.0, "p": 0.0, "p_mod_const": 0.0}
    for k in sample:
        c = cases[k]
        s = iff.Case(c["xf"], c["yf"], precision=iff.IF_FP64, dt=c["dt"], re=c["re"], bodies=c["bodies"],
                     ppe_tol=1e-6, check_every=a.fp64_check_every)
        s.graphs(not a.no_graphs)
        s.persistent(not a.no_persistent)
        s.step(1)
        t0 = time.perf_counter()
        s.step(a.batch_steps)
        tot += time.perf_counter() - t0
        e = fields_err(lambda f: b.get(slots[k], f), s.get)
        err = {f: max(err[f], e[f]) for f in err}
        s.close()
    b.close()
    single_cps = len(sample) / (200.0 * tot / a.batch_steps)
    log("batched: %.3f s/step for 256 cases, %.2f cases/s (FP64 single-case %.3f cases/s)"
        % (per_step, cps, single_cps))
    return dict(value=round(cps, 3), unit="cases/s", cases=len(cases), steps_timed=a.batch_steps,
                ms_per_batched_step=round(1e3 * per_step, 2), mcell_steps_per_s=round(cells / per_step / 1e6, 2),
                sweep_gbs=round(gbs, 1) if gbs else None,
                sweep_roofline_frac=round(gbs / HBM_PEAK_GBS, 3) if gbs else None, max_div=max(divs),
                fp64_single_case_cases_per_s=round(single_cps, 4),
                error_vs_fp64={f: float("%.3g" % v) for f, v in err.items()},
                speedup_vs_fp64_single=round(cps / single_cps, 2), setup_s=round(t_setup, 2),
                autotune=dict(kernel=kernel_name(tr["kernel"], tr["tile_y"], tr["sweeps_per_launch"], "float"),
                              check_every=tr["check_every"], sweep_ns_per_cell=round(tr["sweep_ns"], 4),
                              check_in_sweeps=round(tr["check_sweeps"], 2), typical_sweeps=tr["typical_sweeps"],
                              candidates=tr["candidates"], seconds=round(t_tune, 3)),
                sample="all 256 cases resident; step 1 untimed, then autotune, then steps 2..%d of 200 timed; "
                       "FP64 single-case (library defaults) on 8 of them (every 32nd), same steps; rates per "
                       "200-step case; error_vs_fp64: max over those 8 of max|mixed - FP64| / max(1, max|FP64|) "
                       "after step %d" % (1 + a.batch_steps, 1 + a.batch_steps))


def complex_c3(iff, configs, a):
    """The complex-geometry workload (4096^2, eight star bodies, Re 1000),
    mixed precision: step 1 untimed, autotune, then --c3-steps timed; the
    FP64 path (library defaults) runs the same steps for the error."""
    cfg = configs.c3()
    NX, NY = len(cfg["xf"]) - 1, len(cfg["yf"]) - 1
    kw = dict(dt=cfg["dt"], re=cfg["re"], bodies=cfg["bodies"], ad_itermax=25, ad_tol=1e-6,
              ppe_itermax=400000, ppe_tol=cfg["ppe_tol"], check_every=a.check_every)
    c = iff.Case(cfg["xf"], cfg["yf"], precision=iff.IF_MIXED, **kw)
    c.graphs(not a.no_graphs)
    c.persistent(not a.no_persistent)
    c.step(1)
    tr = c.autotune()
    t0 = time.perf_counter()
    st = c.step(a.c3_steps)
    dt = time.perf_counter() - t0
    c64 = iff.Case(cfg["xf"], cfg["yf"], precision=iff.IF_FP64, **dict(kw, check_every=a.fp64_check_every))
    t1 = time.perf_counter()
    s64 = c64.step(1 + a.c3_steps)
    dt64 = time.perf_counter() - t1
    err = fields_err(c.get, c64.get)
    c.close()
    c64.close()
    log("c3: %.3f s/step mixed, %.3f s/step FP64" % (dt / a.c3_steps, dt64 / (1 + a.c3_steps)))
    return dict(value=round(NX * NY * a.c3_steps / dt / 1e6, 3), unit="Mcell-steps/s", NX=NX, NY=NY,
                ms_per_step=round(1e3 * dt / a.c3_steps, 2), sweeps_last_step=st["ppe_sweeps"],
                max_div=st["max_div"], ppe_tol=cfg["ppe_tol"],
                fp64_mcell_steps_per_s=round(NX * NY * (1 + a.c3_steps) / dt64 / 1e6, 3),
                fp64_max_div=s64["max_div"],
                error_vs_fp64={f: float("%.3g" % v) for f, v in err.items()},
                autotune=dict(kernel=kernel_name(tr["kernel"], tr["tile_y"], tr["sweeps_per_launch"], "float"),
                              check_every=tr["check_every"]),
                sample="step 1 untimed, autotune, steps 2..%d timed; FP64 from the same start, %d steps "
                       "(the FP64 rate includes its first step)" % (1 + a.c3_steps, 1 + a.c3_steps))


def line_sor_f2(iff, configs, a):
    """The deck's PPE_Solver 2 (zebra line SOR, one Thomas solve per row):
    whole steps of the 256^2 cylinder to its 1e-8 stop, FP64 and mixed; and
    the time per sweep on the 2048x1024 workload, from the sweep events of a
    solve capped at 400 sweeps (a full solve there takes ~10^4 sweeps)."""
    out = {}
    runs = [("c1", configs.c1, 2, 200000), ("c2", configs.c2, 1, 400)]
    for cname, mk, steps, cap in (r for r in runs if r[0] in a.line_cases.split(",")):
        cfg = mk()
        NX, NY = len(cfg["xf"]) - 1, len(cfg["yf"]) - 1
        for pname, prec in (("fp64", iff.IF_FP64), ("mixed", iff.IF_MIXED)):
            c = iff.Case(cfg["xf"], cfg["yf"], precision=prec, dt=cfg["dt"], re=cfg["re"], bodies=cfg["bodies"],
                         ad_itermax=25, ad_tol=1e-6, ppe_solver=iff.IF_PPE_LINE, ppe_itermax=cap,
                         ppe_tol=cfg["ppe_tol"], check_every=40)
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
                     "sweep chunks (both colours, forward elimination and back substitution)")
    return out


def serving_c5(iff, configs, a):
    """A bounded sample of the serving workload: the seeded Poisson-arrival
    trace (50 req/s) with grids clipped to <= 512^2 and step counts / 20."""
    from immerseflow_amd import serve
    trace = []
    for t, c in configs.c5_trace(a.serve_requests):
        n = min(len(c["xf"]) - 1, 512)
        xf = configs.uniform_faces(n, 20.0)
        c = dict(c, xf=xf, yf=xf.copy(), dt=0.5 * 20.0 / n, steps=max(5, c["steps"] // 20), ppe_tol=1e-6,
                 check_every=a.check_every)
        trace.append((t, c))
    sample = set(range(0, len(trace), 8))
    recs, summ = serve.serve(trace, precision=iff.IF_MIXED, hbm_budget=16 << 30, tune="auto", keep_fields=sample)
    log("serving:", summ)
    # the error of sampled results against the FP64 path (after the loop)
    err = {"u": 0.0, "v": 0.0, "p": 0.0, "p_mod_const": 0.0}
    for r in recs:
        if r["id"] in sample and r["state"] == 1:
            e = serve.error_vs_fp64(r, trace[r["id"]][1])
            err = {f: max(err[f], e[f]) for f in err}
    return dict(value=round(summ["cases_per_s"], 3), unit="cases/s", requests=summ["requests"],
                completed=summ["completed"], p50_s=round(summ["p50_s"], 4), p99_s=round(summ["p99_s"], 4),
                mcell_steps_per_s=round(summ["mcell_steps_per_s"], 2), max_div=summ["max_div"],
                max_resident=summ["max_resident"], budget_bytes=16 << 30,
                error_vs_fp64={f: float("%.3g" % v) for f, v in err.items()},
                sample="first %d requests of the seeded trace (mean 50 req/s), grids clipped to <= 512^2, "
                       "steps/20 (5..50), 20%% heaving cylinders; error_vs_fp64: max over every 8th request of "
                       "max|served - FP64 alone| / max(1, max|FP64|)" % a.serve_requests)


def gpu_omega(cfg):
    """The library's default relaxation: 2/(1 + sin(pi/N)), N the larger cell count."""
    import math
    n = max(len(cfg["xf"]) - 1, len(cfg["yf"]) - 1)
    return 2.0 / (1.0 + math.sin(math.pi / n))


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
    cfg = configs.c2()
    NX, NY = len(cfg["xf"]) - 1, len(cfg["yf"]) - 1
    cells = NX * NY
    log("workload", cfg["name"], "ratios", ["%.6f" % r for r in cfg["ratios"]])

    case = make_case(iff, cfg, iff.IF_MIXED, a)
    mixed = timed_run(case, a.warmup, a.steps, cells, "mixed")
    sweep_bpc, step_bpc = case.byte_model()
    sweeps_per_launch = int(round(sweep_bpc / 13.0))  # 13 B per cell-sweep (mixed)
    pers = sweeps_per_launch > a.sweeps_per_launch    # a whole check interval per launch
    avg_ms = mixed["sweep_ms"] / max(1, mixed["sweep_launches"])
    achieved = sweep_bpc * cells / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    traffic = None
    tp = os.path.join(ROOT, "profiles", "traffic_rbsor_reg_c2.json")
    kname = kernel_name(a.kernel, a.tile_y, a.sweeps_per_launch, "float", pers)
    if os.path.exists(tp):
        with open(tp) as f:
            tj = json.load(f)
        # measured for this launch shape and sweeps per launch only
        if tj.get("kernel") == kname and tj.get("sweeps_per_launch") == sweeps_per_launch:
            traffic = tj.get("mixed_bytes_per_launch")
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                kernel=kname,
                bytes_per_launch=sweep_bpc * cells, sweeps_per_launch=sweeps_per_launch,
                mean_launch_us=round(avg_ms * 1e3, 3), launches_timed=mixed["sweep_launches"])

    fp64 = None
    if not a.no_fp64 and a.fp64_steps > 0:
        c64 = make_case(iff, cfg, iff.IF_FP64, a)
        r64 = timed_run(c64, max(1, a.warmup), a.fp64_steps, cells, "fp64")
        b64, _ = c64.byte_model()
        avg64 = r64["sweep_ms"] / max(1, r64["sweep_launches"])
        # the mixed path's error: a fresh mixed case run the same steps
        nst = max(1, a.warmup) + a.fp64_steps
        cm = make_case(iff, cfg, iff.IF_MIXED, a)
        cm.step(nst)
        err = fields_err(cm.get, c64.get)
        cm.close()
        # and at a matched tight Poisson stop (1e-9), where the stopping
        # error of either solve no longer shows: the precision effect alone
        tight = dict(cfg, ppe_tol=1e-9)
        ct = [make_case(iff, tight, pr, a) for pr in (iff.IF_MIXED, iff.IF_FP64)]
        for x in ct:
            x.step(3)
        err_tight = fields_err(ct[0].get, ct[1].get)
        for x in ct:
            x.close()
        pers64 = int(round(b64 / 25.0)) > a.fp64_sweeps_per_launch  # 25 B per cell-sweep (FP64)
        fp64 = dict(value=round(r64["mcs"], 3), ms_per_step=round(1e3 * r64["seconds"] / a.fp64_steps, 3),
                    kernel=kernel_name(a.fp64_kernel, a.fp64_tile_y, a.fp64_sweeps_per_launch, "double", pers64),
                    us_per_sweep=round(1e3 * avg64 / max(1, int(round(b64 / 25.0))), 3),
                    sweeps_last_step=r64["stats"]["ppe_sweeps"],
                    max_div=r64["stats"]["max_div"],
                    sweep_gbs=round(b64 * cells / (avg64 * 1e-3) / 1e9, 1) if avg64 > 0 else None,
                    mixed_error_vs_fp64={f: float("%.3g" % v) for f, v in err.items()},
                    error_after_steps=nst,
                    mixed_error_vs_fp64_at_1e9={f: float("%.3g" % v) for f, v in err_tight.items()},
                    error_note="max|mixed - FP64| / max(1, max|FP64|), same case and steps; at the production "
                               "stop (max|div| <= 1e-6) the two solves stop at different iterates, so p differs "
                               "by the stopping error; at a 1e-9 stop (3 steps) only precision differs")
        c64.close()

    # the FP32-velocity variant of the mixed path (IF_MIXED_V32), same case
    v32 = None
    if not a.no_v32:
        cv = make_case(iff, cfg, iff.IF_MIXED_V32, a)
        rv = timed_run(cv, a.warmup, a.steps, cells, "mixed_v32")
        v32 = dict(value=round(rv["mcs"], 3), ms_per_step=round(1e3 * rv["seconds"] / a.steps, 3),
                   max_div=rv["stats"]["max_div"])
        cv.close()
        if fp64 is not None:  # error after the FP64 run's step count, as for the mixed path
            cv = make_case(iff, cfg, iff.IF_MIXED_V32, a)
            cv.step(fp64["error_after_steps"])
            c64 = make_case(iff, cfg, iff.IF_FP64, a)
            c64.step(fp64["error_after_steps"])
            v32["error_vs_fp64"] = {f: float("%.3g" % v) for f, v in fields_err(cv.get, c64.get).items()}
            cv.close()
            c64.close()

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(cfg, case, mixed["stats"], a.omega if a.omega > 0 else gpu_omega(cfg))
        log("cpu baseline", cpu)

    batched = None if a.no_batch else batched_c4(iff, configs, a)
    c3 = None if a.no_c3 else complex_c3(iff, configs, a)
    serving = None if a.no_serve else serving_c5(iff, configs, a)
    line = None if a.no_line else line_sor_f2(iff, configs, a)

    st = mixed["stats"]
    out = {
        "metric": "Mcell-steps/s + % HBM roofline; batched cases/s and p50/p99 latency at div<=1e-6",
        "value": round(mixed["mcs"], 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * mixed["seconds"] / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": cfg["name"], "NX": NX, "NY": NY, "dt": cfg["dt"], "Re": cfg["re"],
                   "precision": "mixed (FP32 Poisson error sweeps in registers, FP64 residual/defect "
                                 "correction, FP64 velocity passes)",
                   "poisson": "red-black SOR, stop max|div|<=%g" % cfg["ppe_tol"],
                   "sweeps_last_step": st["ppe_sweeps"], "ad_iters_last_step": st["ad_iters"],
                   "max_div": st["max_div"], "parallelism": "single GPU",
                   "poisson_launch": ("persistent launch per check interval" if pers else
                                      "launch per %d sweeps" % a.sweeps_per_launch) +
                                     ("" if a.no_graphs else ", hipGraph replay")},
        "roofline": roof,
        "cpu_baseline": cpu,
        "fp64": fp64,
        "mixed_fp32_velocity": v32,
        "batched": batched,
        "complex": c3,
        "serving": serving,
        "line_sor": line,
        "step_bytes_per_cell": round(step_bpc, 1),
    }
    case.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
