"""ctypes binding of the C oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: it is the checker the product is compared
against, never part of the product.  This module builds it on first use when
the shared library is missing (gcc + OpenMP, a few seconds).
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

OF_FLUID, OF_GHOST = 1, 2
OF_OPEN_E, OF_OPEN_W, OF_OPEN_N, OF_OPEN_S = 4, 8, 16, 32
CIRCLE, ELLIPSE, POLYGON = 0, 1, 2

_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_i8p = C.POINTER(C.c_int8)
_i32p = C.POINTER(C.c_int32)


class Body(C.Structure):
    _fields_ = [("kind", C.c_int), ("xc", C.c_double), ("yc", C.c_double),
                ("r", C.c_double), ("a", C.c_double), ("b", C.c_double),
                ("c", C.c_double), ("s", C.c_double), ("nvert", C.c_int),
                ("vx", _dp), ("vy", _dp), ("ub", C.c_double),
                ("vb", C.c_double), ("axb", C.c_double), ("ayb", C.c_double)]


class Ghost(C.Structure):
    _fields_ = [("gidx", C.c_int32), ("body", C.c_int32),
                ("node", C.c_int32 * 4), ("wd", C.c_double * 4),
                ("cd", C.c_double), ("wn", C.c_double * 4),
                ("lnx", C.c_double), ("lny", C.c_double),
                ("bix", C.c_double), ("biy", C.c_double),
                ("ipx", C.c_double), ("ipy", C.c_double)]


class Params(C.Structure):
    _fields_ = [("dt", C.c_double), ("re", C.c_double),
                ("conv_factor", C.c_double), ("ubc", C.c_double),
                ("vbc", C.c_double), ("ad_itermax", C.c_int),
                ("ad_tol", C.c_double), ("ppe_solver", C.c_int),
                ("ppe_omega", C.c_double), ("ppe_itermax", C.c_int),
                ("ppe_tol", C.c_double), ("check_every", C.c_int),
                ("ad_omega", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("ad_iters", C.c_int), ("ad_res", C.c_double),
                ("ppe_sweeps", C.c_int), ("ppe_res", C.c_double),
                ("div_max", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    src = os.path.join(ORACLE_DIR, "oracle.c")
    if (not os.path.exists(LIB_PATH)
            or os.path.getmtime(LIB_PATH) < os.path.getmtime(src)):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    L = C.CDLL(LIB_PATH)
    L.of_grid_metric.argtypes = [C.c_int, _dp, _dp, _dp]
    L.of_compat_init.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.of_compat_ad_step.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_double,
                                    C.c_double, C.c_int, _dp, _dp, _dp, _dp,
                                    _dp, _dp]
    L.of_compat_ad_step.restype = C.c_int
    L.of_compat_ad_step_w.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_double,
                                      C.c_double, C.c_int, C.c_double, _dp, _dp, _dp,
                                      _dp, _dp, _dp]
    L.of_compat_ad_step_w.restype = C.c_int
    L.of_compat_ppe.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_int, _dp, _dp]
    L.of_compat_ppe.restype = C.c_int
    L.of_classify.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_int,
                              C.POINTER(Body), _u8p, _i8p]
    L.of_ghost_list.argtypes = [C.c_int, C.c_int, _u8p, _i32p, C.c_int]
    L.of_ghost_list.restype = C.c_int
    L.of_ghost_weights.argtypes = [C.c_int, C.c_int, _dp, _dp, _u8p, C.c_int,
                                   C.POINTER(Body), _i8p, C.c_int, _i32p,
                                   C.POINTER(Ghost)]
    L.of_closest_point.argtypes = [C.POINTER(Body), C.c_double, C.c_double,
                                   _dp, _dp]
    phys_sig = [C.c_int, C.c_int, _dp, _dp, _u8p, _i8p, C.c_int, C.POINTER(Body),
                C.c_int, C.POINTER(Ghost), C.POINTER(Params), _dp, _dp, _dp,
                _dp, _dp]
    L.of_phys_move_bodies.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_int, C.POINTER(Body),
                                      _u8p, _i8p, _dp, _dp]
    L.of_phys_move_bodies.restype = C.c_int
    L.of_phys_init.argtypes = phys_sig
    L.of_phys_restart.argtypes = phys_sig
    L.of_phys_step.argtypes = phys_sig + [C.POINTER(Stats)]
    L.of_phys_rhs.argtypes = [C.c_int, C.c_int, _dp, _dp, _u8p, C.c_double,
                              _dp, _dp, _dp]
    L.of_phys_residual.argtypes = [C.c_int, C.c_int, _dp, _dp, _u8p, _dp, _dp, _dp]
    L.of_phys_residual.restype = C.c_double
    L.of_phys_poisson.argtypes = [C.c_int, C.c_int, _dp, _dp, _u8p, _dp,
                                  C.POINTER(Params), _dp, _dp]
    L.of_phys_poisson.restype = C.c_int
    L.of_phys_divergence.argtypes = [C.c_int, C.c_int, _dp, _dp, _u8p, _dp, _dp, _dp]
    L.of_phys_divergence.restype = C.c_double
    L.of_phys_ghost_velocity.argtypes = [C.c_int, C.c_int, C.c_int,
                                         C.POINTER(Ghost), C.c_int,
                                         C.POINTER(Body), _dp, _dp]
    L.of_phys_ghost_pressure.argtypes = [C.c_int, C.c_int, C.c_int,
                                         C.POINTER(Ghost), C.c_int,
                                         C.POINTER(Body), _dp]
    _lib = L
    return L


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def u8ptr(a):
    assert a.dtype == np.uint8 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u8p)


def i8ptr(a):
    assert a.dtype == np.int8 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_i8p)


def i32ptr(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_i32p)


def metric(xf):
    xf = np.ascontiguousarray(xf, dtype=np.float64)
    n = len(xf) + 1
    xc = np.zeros(n)
    dx = np.zeros(n)
    assert lib().of_grid_metric(len(xf), dptr(xf), dptr(xc), dptr(dx)) == 0
    return xc, dx


def read_grid_file(path):
    """Face coordinates of a reference grid file ("<id> <x>" per line)."""
    return np.loadtxt(path, usecols=1, dtype=np.float64)


def format_tecplot(nx, ny, xc, yc, f):
    """The reference writer's text (postSim.cu:41-66) -- '%f' of each value."""
    out = ['TITLE = "Post Processing Tecplot"\n',
           'VARIABLES = "X","Y","T"\n',
           'ZONE T="BIG ZONE", I=%d, J=%d, DATAPACKING=POINT\n' % (nx, ny)]
    for j in range(ny):
        for i in range(nx):
            out.append("%f,%f,%f\n" % (xc[i], yc[j], f[i + j * nx]))
    return "".join(out)


class BodySet:
    """Holds ctypes Body records plus the vertex arrays they point into."""

    def __init__(self, bodies):
        self._keep = []
        self.arr = (Body * max(1, len(bodies)))()
        self.n = len(bodies)
        for k, b in enumerate(bodies):
            rec = self.arr[k]
            rec.kind = b["kind"]
            rec.xc, rec.yc = b.get("xc", 0.0), b.get("yc", 0.0)
            rec.r = b.get("r", 0.0)
            rec.a, rec.b = b.get("a", 0.0), b.get("b", 0.0)
            rec.c, rec.s = b.get("c", 1.0), b.get("s", 0.0)
            rec.ub, rec.vb = b.get("ub", 0.0), b.get("vb", 0.0)
            rec.axb, rec.ayb = b.get("axb", 0.0), b.get("ayb", 0.0)
            if b["kind"] == POLYGON:
                vx = np.ascontiguousarray(b["vx"], dtype=np.float64)
                vy = np.ascontiguousarray(b["vy"], dtype=np.float64)
                self._keep += [vx, vy]
                rec.nvert = len(vx)
                rec.vx, rec.vy = dptr(vx), dptr(vy)


def classify(xc, yc, bodyset):
    nx, ny = len(xc), len(yc)
    flags = np.zeros(nx * ny, np.uint8)
    bid = np.zeros(nx * ny, np.int8)
    lib().of_classify(nx, ny, dptr(xc), dptr(yc), bodyset.n, bodyset.arr,
                      u8ptr(flags), i8ptr(bid))
    return flags, bid


def ghost_records(xc, yc, flags, bid, bodyset):
    nx, ny = len(xc), len(yc)
    cap = nx * ny
    lst = np.zeros(cap, np.int32)
    ng = lib().of_ghost_list(nx, ny, u8ptr(flags), i32ptr(lst), cap)
    lst = np.ascontiguousarray(lst[:ng])
    gh = (Ghost * max(1, ng))()
    lib().of_ghost_weights(nx, ny, dptr(xc), dptr(yc), u8ptr(flags),
                           bodyset.n, bodyset.arr, i8ptr(bid), ng,
                           i32ptr(lst), gh)
    return lst, gh, ng


class PhysCase:
    """An oracle physics-mode case (FP64, reference layouts)."""

    def __init__(self, xf, yf, bodies, prm):
        self.xc, self.dx = metric(xf)
        self.yc, self.dy = metric(yf)
        self.nx, self.ny = len(self.xc), len(self.yc)
        nx, ny = self.nx, self.ny
        self.bodies = BodySet(bodies)
        self.flags, self.bid = classify(self.xc, self.yc, self.bodies)
        self.glist, self.gh, self.ng = ghost_records(
            self.xc, self.yc, self.flags, self.bid, self.bodies)
        self.prm = Params(**prm)
        self.u = np.zeros(nx * ny)
        self.v = np.zeros(nx * ny)
        self.p = np.zeros(nx * ny)
        self.uf = np.zeros((nx - 1) * (ny - 2))
        self.vf = np.zeros((nx - 2) * (ny - 1))
        lib().of_phys_init(nx, ny, dptr(self.dx), dptr(self.dy),
                           u8ptr(self.flags), i8ptr(self.bid), self.bodies.n,
                           self.bodies.arr, self.ng, self.gh, C.byref(self.prm), dptr(self.u),
                           dptr(self.v), dptr(self.uf), dptr(self.vf),
                           dptr(self.p))

    def heave(self, specs, t):
        """Place heaving bodies at time t (y = y0 + A sin(2 pi f t)), reclassify
        and rebuild the ghost records.  specs: {body index: (y0, A, f)}."""
        for k, (y0, A, f) in specs.items():
            w = 2.0 * math.pi * f
            b = self.bodies.arr[k]
            b.yc = y0 + A * math.sin(w * t)
            b.vb = A * w * math.cos(w * t)
            b.ayb = -A * w * w * math.sin(w * t)
        fresh = lib().of_phys_move_bodies(self.nx, self.ny, dptr(self.xc), dptr(self.yc),
                                          self.bodies.n, self.bodies.arr, u8ptr(self.flags),
                                          i8ptr(self.bid), dptr(self.u), dptr(self.v))
        self.glist, self.gh, self.ng = ghost_records(self.xc, self.yc, self.flags, self.bid,
                                                     self.bodies)
        return fresh

    def restart(self, u, v, p):
        """Load u, v, p and rebuild ghost velocities and faces (of_phys_restart)."""
        self.u[:] = u
        self.v[:] = v
        self.p[:] = p
        lib().of_phys_restart(self.nx, self.ny, dptr(self.dx), dptr(self.dy), u8ptr(self.flags),
                              i8ptr(self.bid), self.bodies.n, self.bodies.arr, self.ng, self.gh,
                              C.byref(self.prm), dptr(self.u), dptr(self.v), dptr(self.uf),
                              dptr(self.vf), dptr(self.p))

    def step(self, n=1):
        st = Stats()
        out = []
        for _ in range(n):
            lib().of_phys_step(self.nx, self.ny, dptr(self.dx), dptr(self.dy),
                               u8ptr(self.flags), i8ptr(self.bid),
                               self.bodies.n, self.bodies.arr, self.ng, self.gh,
                               C.byref(self.prm), dptr(self.u), dptr(self.v),
                               dptr(self.uf), dptr(self.vf), dptr(self.p),
                               C.byref(st))
            out.append((st.ad_iters, st.ad_res, st.ppe_sweeps, st.ppe_res,
                        st.div_max))
        return out
