"""ctypes declarations of the C ABI (include/immerseflow/if_abi.h)."""
import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libimmerseflow_amd.so")

IF_OK = 0
IF_ERR_ARG, IF_ERR_IO, IF_ERR_PARSE, IF_ERR_HIP = -1, -2, -3, -4
IF_ERR_NOMEM, IF_ERR_STATE, IF_ERR_NODEV, IF_ERR_BUDGET, IF_ERR_DIVERGED = -5, -6, -7, -8, -9
IF_MODE_PHYSICS, IF_MODE_COMPAT = 0, 1
IF_FP64, IF_MIXED, IF_MIXED_V32 = 0, 1, 2
IF_PPE_JACOBI, IF_PPE_RBSOR, IF_PPE_LINE = 0, 1, 2
IF_BODY_CIRCLE, IF_BODY_ELLIPSE, IF_BODY_POLYGON = 0, 1, 2
IF_MOTION_NONE, IF_MOTION_HEAVE = 0, 1
IF_FIELD_U, IF_FIELD_V, IF_FIELD_P, IF_FIELD_UF, IF_FIELD_VF, IF_FIELD_IBLANK, IF_FIELD_FLAGS = range(7)
IF_MAX_BODIES = 16

_dp = C.POINTER(C.c_double)


class IfBody(C.Structure):
    _fields_ = [("kind", C.c_int), ("xc", C.c_double), ("yc", C.c_double),
                ("r", C.c_double), ("a", C.c_double), ("b", C.c_double),
                ("cth", C.c_double), ("sth", C.c_double), ("nvert", C.c_int),
                ("vx", _dp), ("vy", _dp), ("motion", C.c_int),
                ("amp", C.c_double), ("freq", C.c_double)]


class IfInputs(C.Structure):
    _fields_ = [("restart", C.c_int), ("restart_time", C.c_int),
                ("nx", C.c_int), ("ny", C.c_int), ("nxf", C.c_int), ("nyf", C.c_int),
                ("lx", C.c_double), ("ly", C.c_double),
                ("w_ad", C.c_int), ("w_ppe", C.c_int), ("ad_itermax", C.c_int),
                ("ppe_itermax", C.c_int), ("ad_solver", C.c_int), ("ppe_solver", C.c_int),
                ("error_max", C.c_double), ("tmax", C.c_double), ("dt", C.c_double),
                ("re", C.c_double), ("mu", C.c_double), ("write_interval", C.c_int),
                ("mode", C.c_int), ("precision", C.c_int), ("conv_factor", C.c_double),
                ("ppe_omega", C.c_double), ("ppe_tol", C.c_double), ("check_every", C.c_int),
                ("ad_omega", C.c_double),
                ("nbody", C.c_int), ("bodies", IfBody * IF_MAX_BODIES),
                ("poly_store", ((C.c_double * 512) * 2) * IF_MAX_BODIES)]


class IfCaseDesc(C.Structure):
    _fields_ = [("mode", C.c_int), ("precision", C.c_int), ("nxf", C.c_int), ("nyf", C.c_int),
                ("xf", _dp), ("yf", _dp), ("dt", C.c_double), ("re", C.c_double),
                ("conv_factor", C.c_double), ("ubc", C.c_double), ("vbc", C.c_double),
                ("ad_itermax", C.c_int), ("ad_tol", C.c_double), ("ppe_solver", C.c_int),
                ("ppe_omega", C.c_double), ("ppe_itermax", C.c_int), ("ppe_tol", C.c_double),
                ("check_every", C.c_int), ("t0", C.c_double), ("nbody", C.c_int),
                ("bodies", C.POINTER(IfBody)), ("ad_omega", C.c_double)]


class IfStepStats(C.Structure):
    _fields_ = [("steps", C.c_int), ("ad_iters", C.c_int), ("ad_res", C.c_double),
                ("ppe_sweeps", C.c_int), ("ppe_res", C.c_double), ("div_max", C.c_double),
                ("total_ad_iters", C.c_longlong), ("total_ppe_sweeps", C.c_longlong),
                ("max_div", C.c_double), ("failed", C.c_int), ("persistent_timeouts", C.c_int),
                ("weight_table", C.c_int)]


class IfGhost(C.Structure):
    _fields_ = [("gidx", C.c_int32), ("body", C.c_int32), ("node", C.c_int32 * 4),
                ("wd", C.c_double * 4), ("cd", C.c_double), ("wn", C.c_double * 4),
                ("lnx", C.c_double), ("lny", C.c_double), ("bix", C.c_double),
                ("biy", C.c_double), ("ipx", C.c_double), ("ipy", C.c_double)]


class IfTuneResult(C.Structure):
    _fields_ = [("kernel", C.c_int), ("tile_y", C.c_int), ("sweeps_per_launch", C.c_int),
                ("check_every", C.c_int), ("sweep_ns", C.c_double), ("check_sweeps", C.c_double),
                ("typical_sweeps", C.c_int), ("candidates", C.c_int)]


# (name, restype, argtypes) of every declared entry point
SIGNATURES = [
    ("if_last_error", C.c_char_p, []),
    ("if_device_query", C.c_int, [C.POINTER(C.c_int), C.c_char_p, C.c_size_t,
                                  C.POINTER(C.c_int), C.POINTER(C.c_size_t)]),
    ("if_parse_inputs", C.c_int, [C.c_char_p, C.POINTER(IfInputs)]),
    ("if_read_grid", C.c_int, [C.c_char_p, C.c_int, _dp]),
    ("if_grid_metric", C.c_int, [C.c_int, _dp, _dp, _dp]),
    ("if_write_tecplot", C.c_int, [C.c_char_p, C.c_int, C.c_int, _dp, _dp, _dp]),
    ("if_write_tecplot_digits", C.c_int, [C.c_char_p, C.c_int, C.c_int, _dp, _dp, _dp, C.c_int]),
    ("if_read_tecplot", C.c_int, [C.c_char_p, C.c_int, C.c_int, _dp, _dp, _dp]),
    ("if_case_restart", C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, _dp]),
    ("if_case_create", C.c_int, [C.POINTER(IfCaseDesc), C.POINTER(C.c_void_p)]),
    ("if_case_destroy", None, [C.c_void_p]),
    ("if_case_ad_step", C.c_int, [C.c_void_p, C.POINTER(IfStepStats)]),
    ("if_case_ppe_solve", C.c_int, [C.c_void_p, C.POINTER(IfStepStats)]),
    ("if_case_step", C.c_int, [C.c_void_p, C.c_int, C.POINTER(IfStepStats)]),
    ("if_case_get_field", C.c_int, [C.c_void_p, C.c_int, _dp, C.c_size_t]),
    ("if_case_set_field", C.c_int, [C.c_void_p, C.c_int, _dp, C.c_size_t]),
    ("if_case_grid", C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp]),
    ("if_case_dims", C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("if_case_ghosts", C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(IfGhost),
                                 C.c_size_t, C.POINTER(C.c_size_t)]),
    ("if_case_tune", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    ("if_batch_tune", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    ("if_case_graphs", C.c_int, [C.c_void_p, C.c_int]),
    ("if_batch_graphs", C.c_int, [C.c_void_p, C.c_int]),
    ("if_case_persistent", C.c_int, [C.c_void_p, C.c_int]),
    ("if_batch_persistent", C.c_int, [C.c_void_p, C.c_int]),
    ("if_case_autotune", C.c_int, [C.c_void_p, C.c_void_p]),
    ("if_case_monitor", C.c_int, [C.c_void_p, C.c_int, _dp, C.c_int]),
    ("if_case_monitor_read", C.c_int, [C.c_void_p, _dp, C.c_size_t, C.POINTER(C.c_size_t),
                                       C.POINTER(C.c_size_t)]),
    ("if_case_monitor_sample", C.c_int, [C.c_void_p]),
    ("if_case_monitor_sample_noalloc", C.c_int, [C.c_void_p]),
    ("if_batch_autotune", C.c_int, [C.c_void_p, C.c_void_p]),
    ("if_case_sweep_timing", C.c_int, [C.c_void_p, C.c_int, _dp, C.POINTER(C.c_longlong)]),
    ("if_case_byte_model", C.c_int, [C.c_void_p, _dp, _dp]),
    ("if_batch_create", C.c_int, [C.c_int, C.c_size_t, C.POINTER(C.c_void_p)]),
    ("if_batch_destroy", None, [C.c_void_p]),
    ("if_batch_add", C.c_int, [C.c_void_p, C.POINTER(IfCaseDesc), C.c_int, C.POINTER(C.c_int)]),
    ("if_batch_step", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    ("if_batch_slot_state", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                      C.POINTER(C.c_int), C.POINTER(IfStepStats)]),
    ("if_batch_get_field", C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, C.c_size_t]),
    ("if_batch_remove", C.c_int, [C.c_void_p, C.c_int]),
    ("if_batch_info", C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_size_t),
                                _dp, C.POINTER(C.c_longlong), _dp]),
    ("if_case_footprint", C.c_int, [C.POINTER(IfCaseDesc), C.c_int, C.POINTER(C.c_size_t)]),
]

_lib = None


def load():
    """Load the in-tree HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libimmerseflow_amd.so is not built; run "
                           "`python -m immerseflow_amd.build` (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L
