"""GPU: the time-out path of the persistent Poisson sweep, forced from the
host.  With the neighbour-wait limit set to one tick (read when an engine
is created), waits for a slower neighbouring tile give up at once: the
kernel still ends normally (every wait is bounded and, once the error flag
is set, skipped), the step reports the failure, and the engine falls back
to one launch per stage."""
import os

import pytest

import immerseflow_amd as iff
from immerseflow_amd import configs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,shape", [(iff.IF_MIXED, (4, 96, 8, 240)), (iff.IF_FP64, (4, 80, 4, 160))])
def test_persistent_timeout_reports_and_falls_back(require_gpu, prec, shape):
    """FP32 error sweeps, and FP64 sweeps (where the stopped blocks leave
    the pressure itself part-way: finite, and the next solve starts from
    it)."""
    print("[gpu-test] persistent time-out path", prec, shape, flush=True)
    cfg = configs.c2()
    old = os.environ.get("IFF_PERSISTENT_TIMEOUT_TICKS")
    os.environ["IFF_PERSISTENT_TIMEOUT_TICKS"] = "1"
    try:
        c = iff.Case(cfg["xf"], cfg["yf"], precision=prec, dt=cfg["dt"], re=cfg["re"],
                     bodies=cfg["bodies"], ppe_omega=1.998, ppe_itermax=20000, ppe_tol=1e-6, check_every=shape[3])
    finally:
        if old is None:
            del os.environ["IFF_PERSISTENT_TIMEOUT_TICKS"]
        else:
            os.environ["IFF_PERSISTENT_TIMEOUT_TICKS"] = old
    c.tune(*shape)
    with pytest.raises(iff.ImmerseFlowError, match="timed out"):
        c.step(1)
    st = c.step(1)  # one launch per stage from here on
    assert st["div_max"] <= 1e-6
    c.close()
