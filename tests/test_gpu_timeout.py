"""GPU: the time-out path of the persistent Poisson sweep, forced from the
host.  With the neighbour-wait limit set to one tick (read when an engine
is created), waits for a slower neighbouring tile give up at once: the
kernel still ends normally (every wait is bounded and, once the error flag
is set, skipped), the convergence check ends that solve, and the step
solves again from the pressure the stopped chunk left with one launch per
stage -- converged to the same stop level -- and reports the fallback."""
import os

import pytest

import immerseflow_amd as iff
from immerseflow_amd import configs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,shape", [(iff.IF_MIXED, (4, 96, 8, 240)), (iff.IF_FP64, (4, 80, 4, 160))])
def test_persistent_timeout_resolves_per_stage(require_gpu, prec, shape):
    """FP32 error sweeps, and FP64 sweeps (where the stopped blocks leave
    the pressure itself part-way: finite, and the solve restarts from it)."""
    print("[gpu-test] persistent time-out path", prec, shape, flush=True)
    cfg = configs.c2()
    kw = dict(precision=prec, dt=cfg["dt"], re=cfg["re"], bodies=cfg["bodies"], ppe_omega=1.998,
              ppe_itermax=20000, ppe_tol=1e-6, check_every=shape[3])
    old = os.environ.get("IFF_PERSISTENT_TIMEOUT_TICKS")
    os.environ["IFF_PERSISTENT_TIMEOUT_TICKS"] = "1"
    try:
        c = iff.Case(cfg["xf"], cfg["yf"], **kw)
    finally:
        if old is None:
            del os.environ["IFF_PERSISTENT_TIMEOUT_TICKS"]
        else:
            os.environ["IFF_PERSISTENT_TIMEOUT_TICKS"] = old
    c.tune(*shape)
    st = c.step(1)
    assert st["persistent_timeouts"] == 1
    assert st["failed"] == 0 and st["div_max"] <= 1e-6 and st["ppe_res"] <= 1e-6
    st = c.step(1)  # one launch per stage from here on
    assert st["persistent_timeouts"] == 0 and st["div_max"] <= 1e-6
    # the same two steps without the persistent form: the same stop level
    r = iff.Case(cfg["xf"], cfg["yf"], **kw)
    r.tune(*shape)
    r.persistent(False)
    sr = r.step(2)
    assert sr["div_max"] <= 1e-6 and sr["persistent_timeouts"] == 0
    c.close()
    r.close()
