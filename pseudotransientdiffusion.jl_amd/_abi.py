"""ctypes binding of libptd.so (declarations mirror include/ptd.h).

The library is the only compute path: loading fails loudly when libptd.so is missing, and every call
raises PTDError on a non-zero status.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libptd.so")

ABI_VERSION = 1

# enums (include/ptd.h)
OK, ERR_ARG, ERR_HIP, ERR_NAN, ERR_NOMEM, ERR_STATE, ERR_NODEV = range(7)
LIN, STEP, NONLIN = 0, 1, 2
PLAIN, PERF = 0, 1
FP64, FP32, MIXED = 0, 1, 2
NORM_RMS, NORM_MAX = 0, 1
ARITH_EXACT, ARITH_FAST = 0, 1
F_H, F_HOLD, F_D, F_QX, F_QY, F_QZ, F_QX2, F_QY2, F_QZ2, F_RESH, F_DTAU = range(11)
NOALLOC_MAX_PARTIALS = 131072   # PTD_NOALLOC_MAX_PARTIALS

STATUS_NAMES = {OK: "PTD_OK", ERR_ARG: "PTD_ERR_ARG", ERR_HIP: "PTD_ERR_HIP", ERR_NAN: "PTD_ERR_NAN",
                ERR_NOMEM: "PTD_ERR_NOMEM", ERR_STATE: "PTD_ERR_STATE", ERR_NODEV: "PTD_ERR_NODEV"}

# every exported symbol of include/ptd.h
EXPORTS = (
    "ptd_abi_version", "ptd_last_error", "ptd_device_check",
    "ptd_ctx_create", "ptd_ctx_destroy", "ptd_ctx_set_params", "ptd_ctx_device_bytes", "ptd_ctx_stream",
    "ptd_ctx_outer_iterations",
    "ptd_set_field", "ptd_get_field", "ptd_field_count",
    "ptd_compute_iter_params", "ptd_compute_flux", "ptd_compute_update", "ptd_compute_flux_res",
    "ptd_check_res", "ptd_assign",
    "ptd_iterate", "ptd_iterate_timed", "ptd_ctx_iterations_per_launch", "ptd_pt_solve", "ptd_solve",
    "ptd_grad_solve", "ptd_pt_solve_noalloc",
    "ptd_batch_create", "ptd_batch_destroy", "ptd_batch_size", "ptd_batch_device_bytes",
    "ptd_batch_set_field", "ptd_batch_get_field", "ptd_batch_solve", "ptd_batch_tiling", "ptd_batch_group_times",
)


class PTDError(RuntimeError):
    def __init__(self, status, message):
        self.status = status
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")


class Params(ctypes.Structure):
    """struct ptd_params."""
    _fields_ = [
        ("ndim", ctypes.c_int32), ("dcase", ctypes.c_int32), ("form", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("nx", ctypes.c_int64), ("ny", ctypes.c_int64), ("nz", ctypes.c_int64),
        ("itmax", ctypes.c_int64), ("nout", ctypes.c_int32), ("stop_norm", ctypes.c_int32),
        ("arith", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("dt", ctypes.c_double), ("vpdt", ctypes.c_double), ("max_l", ctypes.c_double),
        ("max_l2", ctypes.c_double), ("resc", ctypes.c_double), ("d0", ctypes.c_double),
        ("re", ctypes.c_double), ("theta_r", ctypes.c_double), ("dtau_rho", ctypes.c_double),
        ("tol", ctypes.c_double), ("norm_len", ctypes.c_double),
    ]


_lib = None


def load():
    """Load libptd.so (once). Raises if it is missing: there is no fallback path."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libptd.so not built: {LIB_PATH} is missing (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    i32, i64, dbl, vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    i64p = ctypes.POINTER(ctypes.c_int64)
    pp = ctypes.POINTER(Params)
    sig = {
        "ptd_abi_version": (i32, []),
        "ptd_last_error": (ctypes.c_char_p, []),
        "ptd_device_check": (i32, [i32, ctypes.c_char_p, i32]),
        "ptd_ctx_create": (i32, [pp, i32, ctypes.POINTER(vp)]),
        "ptd_ctx_destroy": (i32, [vp]),
        "ptd_ctx_set_params": (i32, [vp, pp]),
        "ptd_ctx_device_bytes": (i64, [vp]),
        "ptd_ctx_outer_iterations": (i64, [vp]),
        "ptd_ctx_stream": (vp, [vp]),
        "ptd_set_field": (i32, [vp, i32, dp, i64]),
        "ptd_get_field": (i32, [vp, i32, dp, i64]),
        "ptd_field_count": (i64, [vp, i32]),
        "ptd_compute_iter_params": (i32, [vp]),
        "ptd_compute_flux": (i32, [vp]),
        "ptd_compute_update": (i32, [vp]),
        "ptd_compute_flux_res": (i32, [vp]),
        "ptd_check_res": (i32, [vp, dp, dp]),
        "ptd_assign": (i32, [vp]),
        "ptd_iterate": (i32, [vp, i64]),
        "ptd_iterate_timed": (i32, [vp, i64, dp]),
        "ptd_ctx_iterations_per_launch": (i32, [vp, ctypes.POINTER(ctypes.c_int32)]),
        "ptd_pt_solve": (i32, [vp, i64p, dp, dp]),
        "ptd_solve": (i32, [vp, i32, i64p, dp]),
        "ptd_grad_solve": (i32, [vp, i32, dp, i64, dp, dp, dp, i64p, i64p]),
        "ptd_pt_solve_noalloc": (i32, [pp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                       ctypes.POINTER(i32), i64p, dp, dp]),
        "ptd_batch_create": (i32, [pp, i32, i32, ctypes.POINTER(vp)]),
        "ptd_batch_destroy": (i32, [vp]),
        "ptd_batch_size": (i32, [vp]),
        "ptd_batch_device_bytes": (i64, [vp]),
        "ptd_batch_set_field": (i32, [vp, i32, i32, dp, i64]),
        "ptd_batch_get_field": (i32, [vp, i32, i32, dp, i64]),
        "ptd_batch_solve": (i32, [vp, i32, i64p, dp, dp, ctypes.POINTER(i32), dp]),
        "ptd_batch_tiling": (i32, [vp, ctypes.POINTER(i32), ctypes.POINTER(i32), dp, i32]),
        "ptd_batch_group_times": (i32, [vp, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32),
                                        ctypes.POINTER(i32), dp, i32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if L.ptd_abi_version() != ABI_VERSION:
        raise RuntimeError(f"libptd ABI {L.ptd_abi_version()} != {ABI_VERSION}")
    _lib = L
    return L


def check(status):
    if status != OK:
        msg = load().ptd_last_error()
        raise PTDError(status, msg.decode() if msg else "")
    return status
