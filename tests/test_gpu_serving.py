"""GPU: the continuous-batching request loop.  Requests that arrive while
others are resident are admitted between batched steps as memory allows;
each served result equals the same case run alone, bit for bit."""
import numpy as np
import pytest

import immerseflow_amd as iff
from immerseflow_amd import serve

pytestmark = pytest.mark.gpu


def _trace():
    reqs = []
    specs = [(48, [iff.circle(5.0, 8.0, 0.8)], 5), (64, [iff.heaving_circle(6.0, 10.0, 0.6, 0.5, 0.2)], 4),
             (96, [iff.ellipse(8.0, 8.0, 1.0, 0.3, 0.4)], 6), (40, [iff.circle(10.0, 10.0, 1.2)], 3),
             (72, [iff.star(9.0, 9.0, 1.5, [0.05, 0.03, 0.02, 0.06, 0.01], [0.1, 1.2, 2.3, 0.4, 5.0])], 5)]
    for k, (n, bodies, steps) in enumerate(specs):
        xf = np.linspace(0.0, 20.0, n + 1)
        reqs.append((0.004 * k, dict(xf=xf, yf=xf.copy(), bodies=bodies, dt=0.5 * 20.0 / n,
                                     re=200.0, steps=steps, ppe_tol=1e-7, check_every=20)))
    return reqs


def test_served_results_equal_alone(require_gpu):
    print("[gpu-test] serving", flush=True)
    trace = _trace()
    need = [iff.Batch.footprint(c["xf"], c["yf"], precision=iff.IF_MIXED, dt=c["dt"], re=c["re"],
                                bodies=c["bodies"]) for _, c in trace]
    # a budget that holds about three cases at once forces queueing
    records, summary = serve.serve(trace, precision=iff.IF_MIXED, hbm_budget=int(sum(sorted(need)[-3:]) * 1.1),
                                   tune=(0, 16, 2, 20), keep_fields=True)
    assert summary["completed"] == len(trace)
    assert summary["max_div"] <= 1e-6
    assert summary["p99_s"] >= summary["p50_s"] > 0
    assert any(r["queue"] > 0 for r in records) or summary["max_resident"] < len(trace)
    for r in records:
        _, c = trace[r["id"]]
        b = iff.Batch(iff.IF_MIXED)
        b.tune(0, 16, 2, 20)
        s = b.add(c["xf"], c["yf"], c["steps"], dt=c["dt"], re=c["re"], bodies=c["bodies"],
                  ppe_tol=c["ppe_tol"], check_every=c["check_every"])
        b.step(c["steps"])
        for f in ("u", "v", "p"):
            assert np.array_equal(b.get(s, f), r[f]), (r["id"], f)
        b.close()


def test_served_results_state_their_fp64_error(require_gpu):
    """Sampled served results against the FP64 path run alone for the same
    steps: within the mixed path's 1e-4 bound, with fields kept only for the
    requested ids."""
    print("[gpu-test] serving error vs fp64", flush=True)
    trace = _trace()
    keep = {0, 2, 4}
    records, summary = serve.serve(trace, precision=iff.IF_MIXED, tune="auto", keep_fields=keep)
    assert summary["completed"] == len(trace)
    for r in records:
        assert ("u" in r) == (r["id"] in keep)
        if r["id"] in keep:
            e = serve.error_vs_fp64(r, trace[r["id"]][1])
            assert max(e["u"], e["v"], e["p"]) <= 1e-4, (r["id"], e)


def test_serving_trace_prefix_unclipped(require_gpu):
    """The serving workload as configured: the first 6 requests of the
    seeded trace (grids 70^2..1948^2, 101..867 steps, star bodies and a
    heaving cylinder, arrivals at 50 req/s), unclipped.  Every request
    completes with max|div| <= its 1e-7 stop level; the two smallest
    served results equal the same requests run alone (bit for bit) and
    are within the mixed path's 1e-4 of the FP64 path run alone."""
    from immerseflow_amd import configs
    print("[gpu-test] serving trace prefix", flush=True)
    trace = [(t, dict(c, check_every=160)) for t, c in configs.c5_trace(6)]
    sizes = [len(c["xf"]) - 1 for _, c in trace]
    assert max(sizes) > 1900 and min(sizes) < 100
    small = sorted(range(len(trace)), key=lambda k: sizes[k] ** 2 * trace[k][1]["steps"])[:2]
    records, summary = serve.serve(trace, precision=iff.IF_MIXED, hbm_budget=64 << 30, tune=(4, 64, 8, 160),
                                   keep_fields=set(small),
                                   log=lambda m: print("[gpu-test]", m, flush=True))
    assert summary["completed"] == len(trace), [(r["id"], r["state"]) for r in records]
    assert summary["max_div"] <= 1e-7
    assert summary["p99_s"] >= summary["p50_s"] > 0
    for r in records:
        if r["id"] not in small:
            continue
        _, c = trace[r["id"]]
        b = iff.Batch(iff.IF_MIXED)
        b.tune(4, 64, 8, 160)
        s = b.add(c["xf"], c["yf"], c["steps"], dt=c["dt"], re=c["re"], bodies=c["bodies"], ppe_tol=c["ppe_tol"],
                  check_every=160)
        b.step(c["steps"])
        for f in ("u", "v", "p"):
            assert np.array_equal(b.get(s, f), r[f]), (r["id"], f)
        b.close()
        e = serve.error_vs_fp64(r, c)
        assert max(e["u"], e["v"], e["p"]) <= 1e-4, (r["id"], e)
