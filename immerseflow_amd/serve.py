"""Continuous-batching request loop for the fractional step.

Requests (a grid, bodies, Re, dt and a step count) arrive on a trace clock.
Between batched steps the loop admits waiting requests into the resident
batch, first come first served, while their HBM footprint fits the budget.
Every batched step advances all resident cases by one fractional step: one
kernel launch per pass covers the tiles of all of them.  A case that reaches
its step count (or fails) leaves the batch at once and frees its memory.
Each request records its arrival, admission and completion times.
"""
import collections
import time

import numpy as np

from . import IF_FP64, IF_MIXED, Batch, Case


def _kwargs(c):
    return dict(dt=c["dt"], re=c["re"], bodies=c.get("bodies", ()), ppe_tol=c.get("ppe_tol", 1e-6),
                ppe_omega=c.get("ppe_omega", 0.0), ad_itermax=c.get("ad_itermax", 25),
                check_every=c.get("check_every", 160), t0=c.get("t0", 0.0))


def _add(batch, c):
    return batch.add(c["xf"], c["yf"], c["steps"], **_kwargs(c))


def pressure_err(pm, pf, flags):
    """max |pm - pf - c| / max(1, max |pf|) over the cells of the Poisson
    operator (fluid with an open face), with c the mean of pm - pf over
    them.  The pressure of the all-Neumann problem is fixed only up to a
    constant there: the iteration never changes that mode and rounding
    moves it differently in each precision, while the solid cells and the
    corners of the ghost ring keep 0 in both, so a mean over all cells does
    not remove it.  (The workloads here have one connected fluid region.)"""
    f = np.asarray(flags).astype(np.int64)
    fluid = ((f & 1) != 0) & ((f & 0x3c) != 0)
    d = np.asarray(pm)[fluid] - np.asarray(pf)[fluid]
    return float(np.max(np.abs(d - d.mean()))) / max(1.0, float(np.max(np.abs(pf))))


def error_vs_fp64(rec, c):
    """The error of a served (mixed-precision) result against the FP64 path:
    runs request c alone in FP64 for the same number of steps and returns
    max |served - FP64| / max(1, max |FP64|) for u, v and p, and for p
    modulo the constant of the fluid region (p_mod_const, pressure_err).
    rec must hold the served fields (serve(..., keep_fields=...))."""
    s = Case(c["xf"], c["yf"], precision=IF_FP64, **_kwargs(c))
    try:
        s.step(rec["steps"])
        out = {}
        rel = lambda a, b: float(np.max(np.abs(a - b))) / max(1.0, float(np.max(np.abs(b))))
        for f in ("u", "v", "p"):
            out[f] = rel(rec[f], s.get(f))
        out["p_mod_const"] = pressure_err(rec["p"], s.get("p"), s.get("flags"))
        return out
    finally:
        s.close()


def serve(trace, precision=IF_MIXED, hbm_budget=64 << 30, tune="auto",
          time_scale=1.0, keep_fields=False, log=None):
    """Run a request trace [(arrival_s, case_dict), ...] to completion.

    tune: (kernel, tile_y, sweeps_per_launch, check_every), or "auto" to run
    the autotuner on the resident batch after its first batched step, and
    again after any batched step whose largest resident grid has more than
    twice the cells of the largest the last tuning saw (the large cases
    take most of the time, and the shape that suits a few small grids
    does not suit them).
    Returns (records, summary).  records: one dict per request with
    arrival / admitted / finished times (s, trace clock), latency, state,
    steps, cells and the final max |div|; with keep_fields (True, or a
    collection of request ids) also the final u, v, p of those requests."""
    batch = Batch(precision, hbm_budget)
    auto = tune == "auto"
    if not auto:
        batch.tune(*tune)
    else:   # until the autotuner has run: the library's measured default shape
        batch.tune(*((3, 32, 4, 160) if precision == IF_FP64 else (4, 64, 8, 160)))
    tuned = None
    tuned_cells = 0   # the largest resident grid at the last tuning
    ntune = 0
    order = sorted(range(len(trace)), key=lambda k: trace[k][0])
    pending = collections.deque()
    active = {}
    records = []
    nxt = 0
    t0 = time.perf_counter()
    batched_steps = 0
    max_resident = 0
    while nxt < len(order) or pending or active:
        now = time.perf_counter() - t0
        while nxt < len(order) and trace[order[nxt]][0] * time_scale <= now:
            k = order[nxt]
            pending.append((k, trace[k][0] * time_scale, trace[k][1]))
            nxt += 1
        while pending:
            k, arr, c = pending[0]
            slot = _add(batch, c)
            if slot is None:
                if not active:
                    raise RuntimeError("request %d does not fit the HBM budget alone" % k)
                break
            pending.popleft()
            active[slot] = (k, arr, time.perf_counter() - t0, c)
        max_resident = max(max_resident, len(active))
        if not active:
            if nxt < len(order):
                time.sleep(max(0.0, trace[order[nxt]][0] * time_scale - (time.perf_counter() - t0)))
            continue
        batch.step(1)
        batched_steps += 1
        if auto:
            running = [sl for sl in active if batch.state(sl)[0] == 0]
            big = max(((len(active[sl][3]["xf"]) - 1) * (len(active[sl][3]["yf"]) - 1) for sl in running), default=0)
            if big > 2 * tuned_cells:
                tuned = batch.autotune()
                tuned_cells = big
                ntune += 1
                if log:
                    log("autotuned (largest grid %d cells): %s" % (big, tuned))
        now = time.perf_counter() - t0
        for slot in list(active):
            state, steps_done, st = batch.state(slot)
            if state != 0:
                k, arr, adm, c = active.pop(slot)
                rec = dict(id=k, arrival=arr, admitted=adm, finished=now, latency=now - arr,
                           queue=adm - arr, state=state, steps=steps_done,
                           cells=(len(c["xf"]) - 1) * (len(c["yf"]) - 1), max_div=st["max_div"])
                if keep_fields is True or (keep_fields and k in keep_fields):
                    rec["u"] = batch.get(slot, "u")
                    rec["v"] = batch.get(slot, "v")
                    rec["p"] = batch.get(slot, "p")
                records.append(rec)
                batch.remove(slot)
                if log:
                    log("request %d done: %d steps, latency %.3f s" % (k, steps_done, rec["latency"]))
    wall = time.perf_counter() - t0
    batch.close()
    lat = np.array([r["latency"] for r in records])
    ok = [r for r in records if r["state"] == 1]
    summary = dict(requests=len(records), completed=len(ok), wall_s=wall,
                   cases_per_s=len(ok) / wall if wall > 0 else 0.0,
                   p50_s=float(np.percentile(lat, 50)) if len(lat) else None,
                   p99_s=float(np.percentile(lat, 99)) if len(lat) else None,
                   mcell_steps_per_s=sum(r["cells"] * r["steps"] for r in ok) / wall / 1e6,
                   max_div=max((r["max_div"] for r in ok), default=None),
                   batched_steps=batched_steps, max_resident=max_resident, autotune=tuned, autotunes=ntune)
    return records, summary
