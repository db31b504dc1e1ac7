"""Host-side library checks (CPU): exported symbols, the deck grammar, grid
files, the Tecplot writer, error paths.  No GPU compute is called."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle_lib as O
from conftest import GOLDEN, ROOT, gpu_present

import immerseflow_amd as iff
from immerseflow_amd import _abi


def _declared_symbols():
    txt = open(os.path.join(ROOT, "include", "immerseflow", "if_abi.h")).read()
    decl = r"^(?:if_status|void|const char\s*\*)\s*\*?\s*(if_[a-z_]+)\s*\("
    return sorted(set(re.findall(decl, txt, re.M)))


def test_library_exports_every_declared_entry_point():
    lib = _abi.load()
    declared = _declared_symbols()
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(n for n, _, _ in _abi.SIGNATURES) == declared


def test_parse_reference_deck():
    d = iff.read_input_file(os.path.join(GOLDEN, "inputs.txt"))
    assert (d["restart"], d["restart_time"]) == (0, 9)
    assert (d["nx"], d["ny"], d["nxf"], d["nyf"]) == (52, 52, 51, 51)
    assert (d["lx"], d["ly"]) == (10.0, 5.0)
    assert (d["w_ad"], d["w_ppe"], d["ad_itermax"], d["ppe_itermax"]) == (1, 1, 25, 100000)
    assert (d["ad_solver"], d["ppe_solver"]) == (1, 1)
    assert (d["error_max"], d["tmax"], d["dt"], d["re"], d["mu"]) == (1e-6, 100.0, 0.001, 150.0, 0.01)
    assert d["write_interval"] == 1000
    assert d["mode"] == iff.IF_MODE_COMPAT and d["bodies"] == []


def test_parse_extension_blocks(tmp_path):
    deck = open(os.path.join(GOLDEN, "inputs.txt")).read() + """
___________________________| Fractional Step |_______________________________
Mode   Precision   Conv-factor   Omega-PPE   Tol-PPE   Check-every
physics  mixed   0.5   1.9   1e-7   20

___________________________| Bodies |_______________________________
Body (circle xc yc r)
circle 5 8 0.5
Body (ellipse xc yc a b angle-deg)
ellipse 10 10 0.5 0.05 10
Body (star xc yc R a2..a6 phi2..phi6 markers)
star 3 3 0.4 0.01 0.02 0.03 0.04 0.05 0 1 2 3 4 400
Body (heave xc yc r amp freq)
heave 5 8 0.5 0.5 0.2
"""
    p = tmp_path / "deck.txt"
    p.write_text(deck)
    d = iff.read_input_file(str(p))
    assert d["mode"] == iff.IF_MODE_PHYSICS and d["precision"] == iff.IF_MIXED
    assert (d["conv_factor"], d["ppe_omega"], d["ppe_tol"], d["check_every"]) == (0.5, 1.9, 1e-7, 20)
    kinds = [b["kind"] for b in d["bodies"]]
    assert kinds == [iff.IF_BODY_CIRCLE, iff.IF_BODY_ELLIPSE, iff.IF_BODY_POLYGON, iff.IF_BODY_CIRCLE]
    e = d["bodies"][1]
    assert abs(e["cth"] - np.cos(np.radians(10))) < 1e-15 and abs(e["sth"] - np.sin(np.radians(10))) < 1e-15
    s = d["bodies"][2]
    assert len(s["vx"]) == 400 and s["r"] >= np.max(np.hypot(s["vx"] - 3, s["vy"] - 3))
    assert d["bodies"][3]["motion"] == iff.IF_MOTION_HEAVE
    # the reference-era keys are unchanged by the extension
    assert (d["nx"], d["ny"], d["ad_itermax"], d["dt"]) == (52, 52, 25, 0.001)
    # the other precision words
    for word, prec in (("fp64", iff.IF_FP64), ("mixed_v32", iff.IF_MIXED_V32)):
        p.write_text(deck.replace("physics  mixed ", "physics  %s " % word))
        assert iff.read_input_file(str(p))["precision"] == prec
    p.write_text(deck.replace("physics  mixed ", "physics  half "))
    with pytest.raises(iff.ImmerseFlowError) as e:
        iff.read_input_file(str(p))
    assert e.value.code == _abi.IF_ERR_PARSE


def test_parse_errors(tmp_path):
    with pytest.raises(iff.ImmerseFlowError) as e:
        iff.read_input_file(str(tmp_path / "missing.txt"))
    assert e.value.code == _abi.IF_ERR_IO
    p = tmp_path / "bad.txt"
    p.write_text("nothing here\n")
    with pytest.raises(iff.ImmerseFlowError) as e:
        iff.read_input_file(str(p))
    assert e.value.code == _abi.IF_ERR_PARSE


def test_grid_reader_and_metric_match_the_oracle():
    for name, n in (("xgrid.dat2", 51), ("xgrid.dat", 181), ("ygrid.dat", 129)):
        xf = iff.read_grid(os.path.join(GOLDEN, name), n)
        assert np.array_equal(xf, O.read_grid_file(os.path.join(GOLDEN, name)))
        xc, dx = iff.grid_metric(xf)
        oxc, odx = O.metric(xf)
        assert np.array_equal(xc, oxc) and np.array_equal(dx, odx)


def test_short_grid_file_is_an_error(tmp_path):
    p = tmp_path / "short.dat"
    p.write_text("1 0.0\n2 0.5\n")
    with pytest.raises(iff.ImmerseFlowError) as e:
        iff.read_grid(str(p), 51)
    assert e.value.code == _abi.IF_ERR_PARSE


def test_writer_reproduces_golden_bytes(tmp_path):
    xf = O.read_grid_file(os.path.join(GOLDEN, "xgrid.dat2"))
    xc, _ = O.metric(xf)
    out = tmp_path / "final_results.dat"
    iff.write_tecplot(str(out), xc, xc, np.ones(52 * 52))
    assert out.read_bytes() == open(os.path.join(GOLDEN, "final_results.dat"), "rb").read()


def test_writer_row_order_matches_reference_kat(tmp_path):
    """test/output.txt (the reference's 3x3 writer known answer): j outer,
    i inner, one row per cell."""
    kat = open(os.path.join(GOLDEN, "writer_kat_output.txt")).read().splitlines()
    out = tmp_path / "k.dat"
    iff.write_tecplot(str(out), np.arange(3.0), np.arange(3.0), np.arange(1.0, 10.0))
    got = out.read_text().splitlines()
    assert got[2] == kat[2]
    for a, b in zip(got[3:], kat[3:]):
        ga, gb = a.split(","), b.split(",")
        assert float(ga[0]) == float(gb[0]) and float(ga[1]) == float(gb[1]) and ga[2] == gb[2]


def test_checkpoint_writer_round_trips_exactly(tmp_path):
    rng = np.random.default_rng(7)
    x, y = np.cumsum(rng.random(5)), np.cumsum(rng.random(4)) - 1.0
    f = rng.standard_normal(20) * 10.0 ** rng.integers(-12, 12, 20)
    f[3], f[7] = 0.0, -0.0
    out = tmp_path / "ck.dat"
    iff.write_tecplot(str(out), x, y, f, digits=17)
    gx, gy, gf = iff.read_tecplot(str(out), 5, 4)
    assert np.array_equal(gf, f) and np.array_equal(gx, x) and np.array_equal(gy, y)
    assert np.signbit(gf[7]) and not np.signbit(gf[3])


def test_reader_reads_the_reference_format(tmp_path):
    """The reference's own p.dat (printed at %f) reads back as its printed values."""
    nx = ny = 52
    x, y, f = iff.read_tecplot(os.path.join(GOLDEN, "p.dat"), nx, ny)
    rows = open(os.path.join(GOLDEN, "p.dat")).read().splitlines()[3:]
    assert len(rows) == nx * ny
    for k in (0, 1, nx, nx * ny - 1, 1234):
        a = [float(t) for t in rows[k].split(",")]
        assert (x[k % nx], y[k // nx], f[k]) == (a[0], a[1], a[2])
    # and the %f writer reproduces the file from what was read
    out = tmp_path / "p.dat"
    iff.write_tecplot(str(out), x, y, f)
    assert out.read_bytes() == open(os.path.join(GOLDEN, "p.dat"), "rb").read()


def test_reader_errors(tmp_path):
    good = tmp_path / "g.dat"
    iff.write_tecplot(str(good), np.arange(3.0), np.arange(2.0), np.arange(6.0))
    with pytest.raises(iff.ImmerseFlowError) as e:
        iff.read_tecplot(str(good), 2, 3)             # zone is 3 x 2
    assert e.value.code == _abi.IF_ERR_PARSE
    short = tmp_path / "s.dat"
    short.write_text("\n".join(good.read_text().splitlines()[:-1]) + "\n")
    with pytest.raises(iff.ImmerseFlowError) as e:
        iff.read_tecplot(str(short), 3, 2)
    assert e.value.code == _abi.IF_ERR_PARSE
    with pytest.raises(iff.ImmerseFlowError) as e:
        iff.read_tecplot(str(tmp_path / "missing.dat"), 3, 2)
    assert e.value.code == _abi.IF_ERR_IO


def test_writer_unwritable_path():
    with pytest.raises(iff.ImmerseFlowError) as e:
        iff.write_tecplot("/nonexistent_dir/x.dat", np.zeros(2), np.zeros(2), np.zeros(4))
    assert e.value.code == _abi.IF_ERR_IO


@pytest.mark.skipif(gpu_present(), reason="checks the no-device error path")
def test_no_device_fails_loudly():
    xf = np.linspace(0, 1, 17)
    with pytest.raises(iff.ImmerseFlowError) as e:
        iff.Case(xf, xf, dt=0.01, re=100.0)
    assert e.value.code == _abi.IF_ERR_NODEV
    with pytest.raises(iff.ImmerseFlowError):
        iff.Batch()


def test_case_descriptor_validation_without_device():
    # bad descriptors are refused before any device work
    lib = _abi.load()
    d, keep = iff.make_desc(np.array([0.0, 1.0]), np.array([0.0, 1.0]), dt=0.1, re=1.0)
    h = C.c_void_p()
    rc = lib.if_case_create(C.byref(d), C.byref(h))
    assert rc in (_abi.IF_ERR_ARG, _abi.IF_ERR_NODEV)
    assert lib.if_case_step(None, 1, None) == _abi.IF_ERR_ARG


def test_footprint_by_precision():
    """Device bytes a case needs (host-side plan, no device): the mixed mode
    adds the FP32 error, residual and SOR-weight fields to the FP64 fields;
    the FP32-velocity mode stores u, v and the sources in 4 bytes."""
    xf = np.linspace(0.0, 16.0, 257)
    fp = {p: iff.Batch.footprint(xf, xf, precision=p, dt=0.01, re=100.0, bodies=[iff.circle(5.0, 8.0, 0.5)])
          for p in (iff.IF_FP64, iff.IF_MIXED, iff.IF_MIXED_V32)}
    assert fp[iff.IF_MIXED] > fp[iff.IF_FP64] > fp[iff.IF_MIXED_V32]
    # four FP32 fields more than FP64 (e twice, r32, w32): rows of 258 cells
    # pitched to 320, each field aligned to 256 bytes
    pitched = 320 * 258
    assert 16 * pitched <= fp[iff.IF_MIXED] - fp[iff.IF_FP64] <= 16 * pitched + 4 * 1024
    with pytest.raises(iff.ImmerseFlowError):
        iff.Batch.footprint(xf, xf, precision=7, dt=0.01, re=100.0)


def test_footprint_of_a_line_solve_case():
    """A zebra line-SOR case holds its row factors for the whole step: per
    row, five per-cell planes of B x 64 and 13 x 64 reduction values in the
    Poisson precision (256 interior cells: B = 4, S = 64 blocks), beyond
    what the same case needs with red-black SOR."""
    xf = np.linspace(0.0, 16.0, 257)
    B, S = 4, 64
    per_row = (5 * B + 13) * 64
    for prec, es in ((iff.IF_FP64, 8), (iff.IF_MIXED, 4)):
        rb, line = (iff.Batch.footprint(xf, xf, precision=prec, dt=0.01, re=100.0, ppe_solver=sv)
                    for sv in (iff.IF_PPE_RBSOR, iff.IF_PPE_LINE))
        want = es * per_row * 258
        assert want <= line - rb <= want + 256, (prec, line - rb)


def test_pressure_error_modulo_the_fluid_constant():
    """serve.pressure_err compares pressures modulo the constant of the
    Poisson operator's cells (fluid with an open face): a constant offset
    there is no error, the never-written ring corners and solid cells (0 in
    both fields) do not count, and a non-constant difference is reported at
    its size relative to max(1, max|p|)."""
    from immerseflow_amd import serve
    ny, nx = 6, 7
    flags = np.zeros((ny, nx), dtype=np.int64)
    flags[1:-1, 1:-1] = 1 | 4 | 8 | 16 | 32     # operator cells
    flags[0, :] = flags[-1, :] = flags[:, 0] = flags[:, -1] = 1  # ring: fluid, no open face
    flags[3, 3] = 2                              # a ghost (solid) cell
    rng = np.random.default_rng(5)
    pf = rng.standard_normal((ny, nx)) * 3.0
    pf[0, 0] = pf[0, -1] = pf[-1, 0] = pf[-1, -1] = 0.0
    pf[3, 3] = 0.0
    pm = pf + 2.5e-5 * (flags & 0x3c != 0)      # the operator cells' constant only
    assert serve.pressure_err(pm.ravel(), pf.ravel(), flags.ravel()) <= 1e-15 * 4
    pm2 = pm.copy()
    pm2[2, 4] += 1e-6 * np.abs(pf).max()         # one cell off the constant
    e = serve.pressure_err(pm2.ravel(), pf.ravel(), flags.ravel())
    assert 0.9e-6 <= e <= 1.0e-6
