"""GPU: the allocation-free monitor sample (if_case_monitor_sample_noalloc)
against if_case_monitor_sample, through the C ABI: the same record bit for
bit, on two calls in a row -- a case, then a smaller one, so the second
call reuses the static staging array the first one filled."""
import ctypes as C

import numpy as np
import pytest

import immerseflow_amd as iff
from immerseflow_amd import _abi

pytestmark = pytest.mark.gpu


def _sample_both(c):
    L = _abi.load()
    rc = L.if_case_monitor_sample(c._h)
    assert rc == 0, L.if_last_error()
    a = c.monitor_read()
    rc = L.if_case_monitor_sample_noalloc(c._h)
    assert rc == 0, L.if_last_error()
    b = c.monitor_read()
    return a, b


def _case(n, body, steps):
    xf = np.linspace(0.0, 8.0, n + 1)
    c = iff.Case(xf, xf, bodies=[body], dt=0.02, re=100.0, ppe_tol=1e-8, check_every=10)
    c.monitor([(5.0, 4.0), (2.0, 6.5)], capacity=8)
    c.step(steps)
    c.monitor_read()
    return c


def test_noalloc_monitor_sample_equals_allocating_one(require_gpu):
    print("[gpu-test] noalloc monitor sample", flush=True)
    big = _case(512, iff.circle(3.0, 4.1, 0.7), 3)
    small = _case(96, iff.ellipse(4.0, 4.0, 1.1, 0.3, 0.4), 2)
    for c in (big, small):
        a, b = _sample_both(c)
        assert a.shape == b.shape == (1, 3 + 3 * 2)
        assert np.array_equal(a, b)
        assert a[0, 1] != 0.0  # a real force
    # sizes beyond the staging array are refused on the host, before any
    # device work: 32768 x 2064 cells are 512 x 129 = 66,048 tiles
    xf = np.linspace(0.0, 64.0, 32768 + 1)
    yf = np.linspace(0.0, 4.0, 2064 + 1)
    huge = iff.Case(xf, yf, bodies=[iff.circle(3.0, 2.0, 0.5)], dt=0.001, re=100.0)
    huge.monitor([], capacity=2)
    L = _abi.load()
    assert L.if_case_monitor_sample_noalloc(huge._h) == _abi.IF_ERR_ARG
    assert b"65,536" in L.if_last_error() or b"65536" in L.if_last_error()
    huge.close()
    # and the array still serves a fitting case afterwards
    a, b = _sample_both(small)
    assert np.array_equal(a, b)
    for c in (big, small):
        c.close()
