"""ImmerseFlow-AMD: the ImmerseFlow++ fractional-step time step on MI355X.

Python host mirror of the reference's solver interface (`struct ImmerseFlow`,
src/header/globalVariables.cuh:70-88) over the C ABI of
libimmerseflow_amd.so.  Device memory, streams and every kernel live in the
HIP library; there is no CPU compute path here.
"""
import ctypes as C
import math

import numpy as np

from . import _abi
from ._abi import (IF_FIELD_FLAGS, IF_FIELD_IBLANK, IF_FIELD_P, IF_FIELD_U,  # noqa: F401
                   IF_FIELD_UF, IF_FIELD_V, IF_FIELD_VF, IF_FP64, IF_MIXED, IF_MIXED_V32,
                   IF_MODE_COMPAT, IF_MODE_PHYSICS, IF_PPE_JACOBI, IF_PPE_RBSOR, IF_PPE_LINE,
                   IF_BODY_CIRCLE, IF_BODY_ELLIPSE, IF_BODY_POLYGON, IF_MOTION_HEAVE)

FIELDS = {"u": IF_FIELD_U, "v": IF_FIELD_V, "p": IF_FIELD_P, "uf": IF_FIELD_UF,
          "vf": IF_FIELD_VF, "iblank": IF_FIELD_IBLANK, "flags": IF_FIELD_FLAGS}


class ImmerseFlowError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (status %d)" % (msg, code))
        self.code = code


def _lib():
    return _abi.load()


def _check(rc, what=""):
    if rc != 0:
        msg = _lib().if_last_error()
        msg = msg.decode() if msg else ""
        raise ImmerseFlowError(rc, (what + ": " if what else "") + msg)


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---------------------------------------------------------------------------
# files: the reference's deck, grid and Tecplot formats
# ---------------------------------------------------------------------------
def read_input_file(path):
    """readInputFile (src/main.cu:10-59): returns the deck as a dict (nx, ny
    include the ghost ring, as after main.cu:55-58)."""
    inp = _abi.IfInputs()
    _check(_lib().if_parse_inputs(path.encode(), C.byref(inp)), "parse " + path)
    out = {k: getattr(inp, k) for k, _ in _abi.IfInputs._fields_
           if k not in ("bodies", "poly_store")}
    out["bodies"] = []
    for k in range(inp.nbody):
        b = inp.bodies[k]
        d = {"kind": b.kind, "xc": b.xc, "yc": b.yc, "r": b.r, "a": b.a, "b": b.b,
             "cth": b.cth, "sth": b.sth, "motion": b.motion, "amp": b.amp, "freq": b.freq}
        if b.kind == IF_BODY_POLYGON:
            d["vx"] = np.array(b.vx[:b.nvert])
            d["vy"] = np.array(b.vy[:b.nvert])
        out["bodies"].append(d)
    return out


def read_grid(path, nf):
    """The file part of readGridData (preSim.cu:267-291)."""
    xf = np.zeros(nf)
    _check(_lib().if_read_grid(path.encode(), nf, _dptr(xf)), "read " + path)
    return xf


def grid_metric(xf):
    """Cell centres (ghost mirrors) and widths of readGridData (preSim.cu:293-355)."""
    xf = np.ascontiguousarray(xf, dtype=np.float64)
    xc = np.zeros(len(xf) + 1)
    dx = np.zeros(len(xf) + 1)
    _check(_lib().if_grid_metric(len(xf), _dptr(xf), _dptr(xc), _dptr(dx)), "metric")
    return xc, dx


def write_tecplot(path, xc, yc, f, digits=0):
    """write_results_to_file (postSim.cu:41-66); digits > 0 writes every
    value at that many significant digits (17: a lossless checkpoint)."""
    xc = np.ascontiguousarray(xc, dtype=np.float64)
    yc = np.ascontiguousarray(yc, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    if digits:
        rc = _lib().if_write_tecplot_digits(path.encode(), len(xc), len(yc), _dptr(xc), _dptr(yc),
                                            _dptr(f), digits)
    else:
        rc = _lib().if_write_tecplot(path.encode(), len(xc), len(yc), _dptr(xc), _dptr(yc), _dptr(f))
    _check(rc, "write " + path)


def read_tecplot(path, nx, ny):
    """Read a file in the writer's format: (x[nx], y[ny], f[nx*ny])."""
    x, y, f = np.zeros(nx), np.zeros(ny), np.zeros(nx * ny)
    _check(_lib().if_read_tecplot(path.encode(), nx, ny, _dptr(x), _dptr(y), _dptr(f)), "read " + path)
    return x, y, f


def face_coords(xc, yc):
    """Coordinates of the uf ((nx-1) x (ny-2)) and vf ((nx-2) x (ny-1))
    arrays in the reference layouts: face midpoints and cell rows."""
    xc, yc = np.asarray(xc, dtype=np.float64), np.asarray(yc, dtype=np.float64)
    return ((0.5 * (xc[:-1] + xc[1:]), yc[1:-1]), (xc[1:-1], 0.5 * (yc[:-1] + yc[1:])))


CHECKPOINT_FILES = ("uc.dat", "vc.dat", "p.dat", "uf.dat", "vf.dat")


def device_query():
    """CUDAQuery (preSim.cu:187-199): (device count, arch, CUs, HBM bytes)."""
    n, cus, mem = C.c_int(), C.c_int(), C.c_size_t()
    name = C.create_string_buffer(64)
    _check(_lib().if_device_query(C.byref(n), name, 64, C.byref(cus), C.byref(mem)), "device")
    return n.value, name.value.decode(), cus.value, mem.value


# ---------------------------------------------------------------------------
# bodies (geometry is plain data; orientation enters as cos/sin)
# ---------------------------------------------------------------------------
def circle(xc, yc, r):
    return {"kind": IF_BODY_CIRCLE, "xc": xc, "yc": yc, "r": r}


def ellipse(xc, yc, a, b, angle_rad):
    return {"kind": IF_BODY_ELLIPSE, "xc": xc, "yc": yc, "a": a, "b": b,
            "cth": math.cos(angle_rad), "sth": math.sin(angle_rad)}


def star(xc, yc, R, amps, phases, n=400):
    """r(t) = R(1 + sum_{k=2..6} a_k cos(k t + phi_k)) sampled at 2 pi m / n."""
    t = 2.0 * np.pi * np.arange(n) / n
    rr = np.full(n, 1.0)
    for k, (a, ph) in enumerate(zip(amps, phases), start=2):
        rr = rr + a * np.cos(k * t + ph)
    rr = R * rr
    vx = xc + rr * np.cos(t)
    vy = yc + rr * np.sin(t)
    return {"kind": IF_BODY_POLYGON, "xc": xc, "yc": yc, "r": float(rr.max()) * 1.0001 + 1e-12,
            "vx": vx, "vy": vy}


def heaving_circle(xc, yc, r, amp, freq):
    d = circle(xc, yc, r)
    d.update(motion=IF_MOTION_HEAVE, amp=amp, freq=freq)
    return d


class _BodyArray:
    def __init__(self, bodies):
        self.keep = []
        self.arr = (_abi.IfBody * max(1, len(bodies)))()
        for k, b in enumerate(bodies):
            r = self.arr[k]
            r.kind = b["kind"]
            r.xc, r.yc, r.r = b.get("xc", 0.0), b.get("yc", 0.0), b.get("r", 0.0)
            r.a, r.b = b.get("a", 0.0), b.get("b", 0.0)
            r.cth, r.sth = b.get("cth", 1.0), b.get("sth", 0.0)
            r.motion, r.amp, r.freq = b.get("motion", 0), b.get("amp", 0.0), b.get("freq", 0.0)
            if b["kind"] == IF_BODY_POLYGON:
                vx = np.ascontiguousarray(b["vx"], dtype=np.float64)
                vy = np.ascontiguousarray(b["vy"], dtype=np.float64)
                self.keep += [vx, vy]
                r.nvert, r.vx, r.vy = len(vx), _dptr(vx), _dptr(vy)


def make_desc(xf, yf, *, mode=IF_MODE_PHYSICS, precision=IF_FP64, dt, re, conv_factor=0.5,
              ubc=1.0, vbc=0.0, ad_itermax=25, ad_tol=1e-6, ppe_solver=IF_PPE_RBSOR,
              ppe_omega=0.0, ppe_itermax=100000, ppe_tol=1e-8, check_every=10, t0=0.0,
              bodies=(), ad_omega=1.0):
    """A case descriptor plus the arrays it points into (keep both alive)."""
    xf = np.ascontiguousarray(xf, dtype=np.float64)
    yf = np.ascontiguousarray(yf, dtype=np.float64)
    ba = _BodyArray(list(bodies))
    d = _abi.IfCaseDesc()
    d.mode, d.precision = mode, precision
    d.nxf, d.nyf, d.xf, d.yf = len(xf), len(yf), _dptr(xf), _dptr(yf)
    d.dt, d.re, d.conv_factor, d.ubc, d.vbc = dt, re, conv_factor, ubc, vbc
    d.ad_itermax, d.ad_tol = ad_itermax, ad_tol
    d.ppe_solver, d.ppe_omega, d.ppe_itermax = ppe_solver, ppe_omega, ppe_itermax
    d.ppe_tol, d.check_every, d.t0 = ppe_tol, check_every, t0
    d.nbody, d.bodies = len(bodies), ba.arr
    d.ad_omega = ad_omega
    return d, (xf, yf, ba)


def _tune_dict(r):
    return {k: getattr(r, k) for k, _ in r._fields_}


def _stats_dict(st):
    return {k: getattr(st, k) for k, _ in _abi.IfStepStats._fields_}


# ---------------------------------------------------------------------------
# one case
# ---------------------------------------------------------------------------
class Case:
    """A resident case (allocation + readGridData + initializeData)."""

    def __init__(self, xf, yf, **kw):
        self._desc, self._keep = make_desc(xf, yf, **kw)
        h = C.c_void_p()
        _check(_lib().if_case_create(C.byref(self._desc), C.byref(h)), "case create")
        self._h = h
        nx, ny = C.c_int(), C.c_int()
        _check(_lib().if_case_dims(self._h, C.byref(nx), C.byref(ny)))
        self.nx, self.ny = nx.value, ny.value
        self.xc, self.yc = np.zeros(self.nx), np.zeros(self.ny)
        self.dx, self.dy = np.zeros(self.nx), np.zeros(self.ny)
        _check(_lib().if_case_grid(self._h, _dptr(self.xc), _dptr(self.yc), _dptr(self.dx),
                                   _dptr(self.dy)))

    def close(self):
        if getattr(self, "_h", None):
            _lib().if_case_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def step(self, n=1):
        """n physics-mode fractional steps; returns the last step's stats."""
        st = _abi.IfStepStats()
        _check(_lib().if_case_step(self._h, n, C.byref(st)), "step")
        return _stats_dict(st)

    def ad_step(self):
        """Compat mode: one ImmerseFlow::ADsolver call."""
        st = _abi.IfStepStats()
        _check(_lib().if_case_ad_step(self._h, C.byref(st)), "ADsolver")
        return _stats_dict(st)

    def ppe_solve(self):
        """Compat mode: ImmerseFlow::PPESolver."""
        st = _abi.IfStepStats()
        _check(_lib().if_case_ppe_solve(self._h, C.byref(st)), "PPESolver")
        return _stats_dict(st)

    def field_size(self, name):
        nx, ny = self.nx, self.ny
        return {"uf": (nx - 1) * (ny - 2), "vf": (nx - 2) * (ny - 1)}.get(name, nx * ny)

    def get(self, name):
        out = np.zeros(self.field_size(name))
        _check(_lib().if_case_get_field(self._h, FIELDS[name], _dptr(out), out.size), "get " + name)
        return out

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        _check(_lib().if_case_set_field(self._h, FIELDS[name], _dptr(a), a.size), "set " + name)

    def ghosts(self):
        """(sorted ghost-cell list, image-point records) in reference indices."""
        n = C.c_size_t()
        _check(_lib().if_case_ghosts(self._h, None, None, 0, C.byref(n)))
        k = n.value
        lst = (C.c_int32 * max(1, k))()
        recs = (_abi.IfGhost * max(1, k))()
        _check(_lib().if_case_ghosts(self._h, lst, recs, k, C.byref(n)))
        return np.array(lst[:k], dtype=np.int32), recs, k

    def tune(self, kernel, tile_y, sweeps_per_launch, check_every):
        """Poisson sweep kernel (0 LDS-tiled, 1 register-streaming, 2/3/4
        register-tile with 4/8/16 waves), tile or strip height, red-black
        sweeps per launch, sweeps per check.  Every kernel and shape gives
        the same result.  Without a call the library uses 4/64/8 (mixed)
        or, in FP64, 4/48, 4/64 or 4/80 with 4 sweeps as persistent
        chunks (the smallest height that fits on the device) else 3/32/4, and
        the case's check_every."""
        _check(_lib().if_case_tune(self._h, kernel, tile_y, sweeps_per_launch, check_every), "tune")

    def restart(self, u, v, p, uf=None, vf=None):
        """Load u, v, p (and the faces; without them the faces are
        interpolated from u, v) -- see if_case_restart."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None for a in (u, v, p, uf, vf)]
        ptrs = [_dptr(a) if a is not None else None for a in arrs]
        _check(_lib().if_case_restart(self._h, *ptrs), "restart")

    def write_checkpoint(self, directory, digits=17):
        """uc.dat, vc.dat, p.dat and the faces uf.dat, vf.dat in the
        writer's format (17 digits: lossless)."""
        import os
        (ux, uy), (vx, vy) = face_coords(self.xc, self.yc)
        for name, arr, xs, ys in (("uc.dat", self.get("u"), self.xc, self.yc),
                                  ("vc.dat", self.get("v"), self.xc, self.yc),
                                  ("p.dat", self.get("p"), self.xc, self.yc),
                                  ("uf.dat", self.get("uf"), ux, uy), ("vf.dat", self.get("vf"), vx, vy)):
            write_tecplot(os.path.join(directory, name), xs, ys, arr, digits=digits)

    def restart_from(self, directory):
        """Restart from write_checkpoint's files, or from the reference's
        uc.dat/vc.dat/p.dat alone (faces then interpolated)."""
        import os
        nx, ny = len(self.xc), len(self.yc)
        u = read_tecplot(os.path.join(directory, "uc.dat"), nx, ny)[2]
        v = read_tecplot(os.path.join(directory, "vc.dat"), nx, ny)[2]
        p = read_tecplot(os.path.join(directory, "p.dat"), nx, ny)[2]
        uf = vf = None
        if os.path.exists(os.path.join(directory, "uf.dat")) and os.path.exists(os.path.join(directory, "vf.dat")):
            uf = read_tecplot(os.path.join(directory, "uf.dat"), nx - 1, ny - 2)[2]
            vf = read_tecplot(os.path.join(directory, "vf.dat"), nx - 2, ny - 1)[2]
        self.restart(u, v, p, uf, vf)

    def monitor(self, probes=(), capacity=65536):
        """Record t, the body force (Fx, Fy) and u, v, p at the cells nearest
        the probe points [(x, y), ...] after every step (if_case_monitor)."""
        xy = np.ascontiguousarray(np.asarray(probes, dtype=np.float64).reshape(-1))
        self._nprobe = len(xy) // 2
        _check(_lib().if_case_monitor(self._h, self._nprobe, _dptr(xy) if len(xy) else None, capacity),
               "monitor")

    def monitor_read(self):
        """Records since the last read: array (n, 3 + 3*nprobe) with columns
        t, Fx, Fy, then u, v, p per probe."""
        n, dropped = C.c_size_t(), C.c_size_t()
        _check(_lib().if_case_monitor_read(self._h, None, 0, C.byref(n), C.byref(dropped)), "monitor")
        out = np.zeros((n.value, 3 + 3 * self._nprobe))
        if n.value:
            _check(_lib().if_case_monitor_read(self._h, _dptr(out), n.value, C.byref(n), C.byref(dropped)),
                   "monitor")
        return out[:n.value]

    def monitor_sample(self):
        """Append one monitor record of the current fields at the current
        time, without a step (if_case_monitor_sample)."""
        _check(_lib().if_case_monitor_sample(self._h), "monitor")

    def autotune(self):
        """Time every Poisson sweep launch shape on the current fields, keep
        the fastest and choose the check interval (see if_case_autotune).
        Returns the choice as a dict; the case's state is unchanged."""
        r = _abi.IfTuneResult()
        _check(_lib().if_case_autotune(self._h, C.byref(r)), "autotune")
        return _tune_dict(r)

    def graphs(self, enable=True):
        """Replay each chunk of Poisson sweeps from a captured hipGraph (the
        default) or launch the kernels one by one; results are identical."""
        _check(_lib().if_case_graphs(self._h, 1 if enable else 0), "graphs")

    def persistent(self, enable=True):
        """Run each chunk of Poisson sweeps as one persistent launch when the
        shape allows it and the grid fits on the device (the default), or
        as one launch per stage; results are identical."""
        _check(_lib().if_case_persistent(self._h, 1 if enable else 0), "persistent")

    def sweep_timing(self, reset=False):
        ms, n = C.c_double(), C.c_longlong()
        _check(_lib().if_case_sweep_timing(self._h, 1 if reset else 0, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def byte_model(self):
        a, b = C.c_double(), C.c_double()
        _check(_lib().if_case_byte_model(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value


# ---------------------------------------------------------------------------
# many cases: batched launches, admission by HBM budget
# ---------------------------------------------------------------------------
class Batch:
    def __init__(self, precision=IF_FP64, hbm_budget=0):
        h = C.c_void_p()
        _check(_lib().if_batch_create(precision, hbm_budget, C.byref(h)), "batch create")
        self._h = h
        self._keep = {}

    def close(self):
        if getattr(self, "_h", None):
            _lib().if_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def footprint(xf, yf, precision=IF_FP64, **kw):
        d, keep = make_desc(xf, yf, precision=precision, **kw)
        n = C.c_size_t()
        _check(_lib().if_case_footprint(C.byref(d), precision, C.byref(n)))
        return n.value

    def add(self, xf, yf, nsteps, **kw):
        """Admit a case; returns its slot, or None when the budget refuses it."""
        d, keep = make_desc(xf, yf, **kw)
        slot = C.c_int()
        rc = _lib().if_batch_add(self._h, C.byref(d), nsteps, C.byref(slot))
        if rc == _abi.IF_ERR_BUDGET:
            return None
        _check(rc, "batch add")
        self._keep[slot.value] = (d, keep, len(xf) + 1, len(yf) + 1)
        return slot.value

    def step(self, n=1):
        fin = C.c_int()
        _check(_lib().if_batch_step(self._h, n, C.byref(fin)), "batch step")
        return fin.value

    def state(self, slot):
        s, k = C.c_int(), C.c_int()
        st = _abi.IfStepStats()
        _check(_lib().if_batch_slot_state(self._h, slot, C.byref(s), C.byref(k), C.byref(st)))
        return s.value, k.value, _stats_dict(st)

    def get(self, slot, name):
        _, _, nx, ny = self._keep[slot]
        n = {"uf": (nx - 1) * (ny - 2), "vf": (nx - 2) * (ny - 1)}.get(name, nx * ny)
        out = np.zeros(n)
        _check(_lib().if_batch_get_field(self._h, slot, FIELDS[name], _dptr(out), n), "get")
        return out

    def remove(self, slot):
        _check(_lib().if_batch_remove(self._h, slot), "remove")
        self._keep.pop(slot, None)

    def tune(self, kernel, tile_y, sweeps_per_launch, check_every):
        """As Case.tune, for every case of the batch."""
        _check(_lib().if_batch_tune(self._h, kernel, tile_y, sweeps_per_launch, check_every), "tune")

    def graphs(self, enable=True):
        """As Case.graphs."""
        _check(_lib().if_batch_graphs(self._h, 1 if enable else 0), "graphs")

    def persistent(self, enable=True):
        """As Case.persistent."""
        _check(_lib().if_batch_persistent(self._h, 1 if enable else 0), "persistent")

    def autotune(self):
        """As Case.autotune, over every running case of the batch."""
        r = _abi.IfTuneResult()
        _check(_lib().if_batch_autotune(self._h, C.byref(r)), "autotune")
        return _tune_dict(r)

    def info(self):
        r, b, ms, n, cs = C.c_int(), C.c_size_t(), C.c_double(), C.c_longlong(), C.c_double()
        _check(_lib().if_batch_info(self._h, C.byref(r), C.byref(b), C.byref(ms), C.byref(n),
                                    C.byref(cs)))
        return {"resident": r.value, "bytes": b.value, "sweep_ms": ms.value, "launches": n.value,
                "cell_sweeps": cs.value}


# ---------------------------------------------------------------------------
# the reference's driver-facing object
# ---------------------------------------------------------------------------
class ImmerseFlow:
    """Mirror of `struct ImmerseFlow` (globalVariables.cuh:70-88) and the
    call order of main() (src/main.cu:75-102)."""

    def __init__(self):
        self.Input = None
        self.case = None
        self.xf = self.yf = None

    def readInputFile(self, path):
        self.Input = read_input_file(path)
        return self.Input

    def CUDAQuery(self):
        return device_query()

    def readGridData(self, xgrid, ygrid):
        self.xf = read_grid(xgrid, self.Input["nxf"])
        self.yf = read_grid(ygrid, self.Input["nyf"])

    def initializeData(self):
        I = self.Input
        kw = dict(mode=I["mode"], precision=I["precision"], dt=I["dt"], re=I["re"],
                  conv_factor=I["conv_factor"], ad_itermax=I["ad_itermax"],
                  ppe_itermax=I["ppe_itermax"], ppe_omega=I["ppe_omega"],
                  ppe_tol=I["ppe_tol"], check_every=I["check_every"],
                  ppe_solver={0: IF_PPE_JACOBI, 2: IF_PPE_LINE}.get(I["ppe_solver"], IF_PPE_RBSOR),
                  ad_tol=1e-6, bodies=I["bodies"],
                  ad_omega=I["ad_omega"] if I["ad_omega"] > 0 else float(I["w_ad"] if I["w_ad"] > 0 else 1))
        self.case = Case(self.xf, self.yf, **kw)
        return self.case

    def ADsolver(self):
        return self.case.ad_step()

    def PPESolver(self):
        return self.case.ppe_solve()

    def step(self, n=1):
        return self.case.step(n)

    def freeAllocation(self):
        if self.case:
            self.case.close()
            self.case = None
