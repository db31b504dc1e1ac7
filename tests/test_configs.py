"""Grid generation and the synthetic workloads (CPU).

The reference ships its grids as files in the format '<1-based id> <x>'
per face (inputs/uniformGrid.py:3-14; read by readGridData,
src/include/preSim.cu:267-291).  configs.py builds the workloads' grids;
these tests pin them against their specification and check that a grid
written in the reference format reads back bit for bit through the
library's reader (the drop-in path of the CLI)."""
import math

import numpy as np

import immerseflow_amd as iff
from immerseflow_amd import configs


def _widths(f):
    return np.diff(f)


def test_c2_grid_spec():
    """2048 x 1024 cells; uniform h = 0.005 on the 4 x 2 box around the body
    (box at x in [13, 17], y in [9, 11] of a 30 x 20 domain); geometric
    stretching outward with one ratio per side, <= 1.02; the domain ends
    exactly at 0 and 30 (20); the ellipse (chord 1, thickness 0.1, 10 deg)
    lies inside the box."""
    cfg = configs.c2()
    xf, yf = cfg["xf"], cfg["yf"]
    assert len(xf) == 2049 and len(yf) == 1025
    assert xf[0] == 0.0 and xf[-1] == 30.0 and yf[0] == 0.0 and yf[-1] == 20.0
    for f, lo, hi, n_box in ((xf, 13.0, 17.0, 800), (yf, 9.0, 11.0, 400)):
        w = _widths(f)
        assert np.all(w > 0)
        inb = (f[:-1] >= lo - 1e-9) & (f[1:] <= hi + 1e-9)
        assert inb.sum() == n_box
        assert np.max(np.abs(w[inb] - 0.005)) < 1e-9
        # outward geometric growth: constant ratio per side
        first = np.argmax(inb)
        last = first + n_box - 1
        left = w[:first][::-1]   # box outward to the west (south)
        right = w[last + 1:]
        for side in (left, right):
            r = side[1:] / side[:-1]
            assert np.all(r > 1.0) and np.all(r <= 1.02)
            assert np.max(r) - np.min(r) < 1e-9
            assert abs(side[0] / 0.005 - r[0]) < 1e-6  # first cell grows from h
    assert all(1.0 < r <= 1.02 for r in cfg["ratios"])
    assert cfg["dt"] == 0.5 * 0.005 and cfg["re"] == 300.0 and cfg["steps"] == 2000
    b = cfg["bodies"][0]
    assert b["kind"] == iff.IF_BODY_ELLIPSE and b["a"] == 0.5 and b["b"] == 0.05
    assert abs(math.atan2(b["sth"], b["cth"]) - math.radians(10.0)) < 1e-15
    assert 13.0 < b["xc"] - b["a"] and b["xc"] + b["a"] < 17.0
    assert 9.0 < b["yc"] - b["a"] and b["yc"] + b["a"] < 11.0


def test_c1_c3_c4_c5_grid_specs():
    c1 = configs.c1()
    assert len(c1["xf"]) == 257 and c1["xf"][-1] == 16.0 and np.allclose(_widths(c1["xf"]), 16.0 / 256)
    assert c1["bodies"][0]["r"] == 0.5 and (c1["bodies"][0]["xc"], c1["bodies"][0]["yc"]) == (5.0, 8.0)
    assert (c1["dt"], c1["re"], c1["steps"], c1["ppe_tol"]) == (0.002, 100.0, 500, 1e-8)
    c3 = configs.c3()
    assert len(c3["xf"]) == 4097 and c3["xf"][-1] == 20.0 and len(c3["bodies"]) == 8
    assert c3["re"] == 1000.0 and c3["steps"] == 1000 and c3["dt"] == 0.5 * 20.0 / 4096
    for b in c3["bodies"]:
        assert len(b["vx"]) == 400
        rr = np.hypot(b["vx"] - b["xc"], b["vy"] - b["yc"])
        assert 0.3 * (1 - 5 * 0.08) <= rr.min() and rr.max() <= 0.6 * (1 + 5 * 0.08) + 1e-12
        assert 2.0 <= b["xc"] - rr.max() and b["xc"] + rr.max() <= 18.0
    # non-overlapping bounding circles
    bs = c3["bodies"]
    for k in range(len(bs)):
        for m in range(k):
            d = math.hypot(bs[k]["xc"] - bs[m]["xc"], bs[k]["yc"] - bs[m]["yc"])
            assert d > bs[k]["r"] / 1.0001 + bs[m]["r"] / 1.0001 - 1e-9
    c4 = configs.c4_cases(256)
    ns = [len(c["xf"]) - 1 for c in c4]
    assert len(c4) == 256 and min(ns) >= 128 and max(ns) <= 1024
    assert all(c["steps"] == 200 and 50 <= c["re"] <= 500 for c in c4)
    # dt: CFL 0.5 capped by the explicit predictor's stability limit
    for c in c4:
        h = 16.0 / (len(c["xf"]) - 1)
        assert c["dt"] == min(0.5 * h, 1.0 / c["re"])
    assert any(c["dt"] < 0.5 * 16.0 / (len(c["xf"]) - 1) for c in c4)
    tr = configs.c5_trace(500)
    ts = [t for t, _ in tr]
    assert len(tr) == 500 and all(b > a for a, b in zip(ts, ts[1:]))
    assert 0.5 * 500 / 50.0 < ts[-1] < 2.0 * 500 / 50.0   # mean inter-arrival 20 ms
    ns = [len(c["xf"]) - 1 for _, c in tr]
    assert min(ns) >= 64 and max(ns) <= 2048
    assert all(100 <= c["steps"] <= 1000 for _, c in tr)
    assert all(c["dt"] == configs.stable_dt(20.0 / (len(c["xf"]) - 1), c["re"]) for _, c in tr)
    heave = sum(1 for _, c in tr if c["bodies"][0].get("motion", 0) == iff.IF_MOTION_HEAVE)
    assert 0.12 * 500 < heave < 0.28 * 500


def test_grid_file_round_trip(tmp_path):
    """write_grid_file -> if_read_grid returns the same doubles (17
    significant digits in the reference's '<id> <x>' format), and the metric
    of the read grid equals the metric of the generated one."""
    cfg = configs.c2()
    for name, f in (("xgrid.dat", cfg["xf"]), ("ygrid.dat", cfg["yf"])):
        p = str(tmp_path / name)
        configs.write_grid_file(p, f)
        back = iff.read_grid(p, len(f))
        assert np.array_equal(back, f)
        xc0, dx0 = iff.grid_metric(f)
        xc1, dx1 = iff.grid_metric(back)
        assert np.array_equal(xc0, xc1) and np.array_equal(dx0, dx1)
        # the file has the reference layout: 1-based ids, one face per line
        with open(p) as fh:
            lines = fh.read().splitlines()
        assert len(lines) == len(f)
        assert lines[0].split()[0] == "1" and lines[-1].split()[0] == str(len(f))
