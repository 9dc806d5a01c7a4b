"""ctypes binding of the libheom_amd C ABI (include/libheom_amd.h).

The shared library is built in-tree into ``libheom_amd/_lib/libheom_amd.so``
by ``libheom_amd.build.build()``.  Loading fails loudly when it is missing:
there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libheom_amd.so")

HEOM_OK = 0
HEOM_ERR_ARG = 1
HEOM_ERR_STATE = 2
HEOM_ERR_HIP = 3
HEOM_ERR_STEP_REJECTED = 4
HEOM_ERR_UNSUPPORTED = 5
HEOM_ERR_NOMEM = 6

HEOM_HILB, HEOM_LIOU, HEOM_ADO = 0, 1, 2
HEOM_REDFIELD_HILB, HEOM_REDFIELD_LIOU = 3, 4
HEOM_C128, HEOM_C64, HEOM_C64_BF16X3 = 0, 1, 2
HEOM_RK4, HEOM_RKDP, HEOM_LSRK4 = 0, 1, 2
HEOM_CLOSURE_PER_ADO, HEOM_CLOSURE_GLOBAL = 0, 1

# every symbol declared in include/libheom_amd.h
EXPORTS = [
    "heom_last_error", "heom_version", "heom_default_options", "heom_create", "heom_destroy",
    "heom_set_H", "heom_alloc_noises", "heom_set_noise", "heom_set_hrchy_filter", "heom_set_param", "heom_get_n_hrchy",
    "heom_get_n_dim", "heom_main_size", "heom_get_hrchy", "heom_redfield_lambda", "heom_calc_diff",
    "heom_calc_diff_host", "heom_batch_create", "heom_batch_layout", "heom_batch_tuning", "heom_batch_set_tuning", "heom_solver_create", "heom_solver_destroy", "heom_solve",
    "heom_solve_dev", "heom_solve_dev_noalloc", "heom_solver_stats", "heom_solver_reset", "heom_solver_set_graphs", "heom_solver_time_stages",
    "heom_solver_stage_bytes", "heom_dev_alloc", "heom_dev_free", "heom_memcpy_h2d",
    "heom_memcpy_d2h", "heom_device_sync", "heom_stream_copy", "heom_mlp_create", "heom_mlp_destroy",
    "heom_mlp_n_params", "heom_mlp_get_params", "heom_mlp_set_params", "heom_mlp_set_scaling",
    "heom_mlp_get_scaling", "heom_mlp_fit_scaling", "heom_mlp_forward", "heom_mlp_loss_grad",
    "heom_mlp_adam_step", "heom_mlp_train", "heom_set_closure", "heom_closure_dims",
    "heom_closure_rows",
]


class HeomError(RuntimeError):
    """A libheom_amd call returned a non-zero status."""

    def __init__(self, code, msg):
        super().__init__("libheom_amd error %d: %s" % (code, msg))
        self.code = code


class heom_options(ctypes.Structure):
    _fields_ = [("parity_q1", ctypes.c_int), ("sqrt_norm", ctypes.c_int),
                ("order", ctypes.c_int), ("device", ctypes.c_int),
                ("matrix_cores", ctypes.c_int), ("renorm_rho0", ctypes.c_int)]


class heom_stats(ctypes.Structure):
    _fields_ = [("n_rhs", ctypes.c_longlong), ("n_attempts", ctypes.c_longlong),
                ("n_accepted", ctypes.c_longlong), ("n_rejected", ctypes.c_longlong),
                ("ado_stage_updates", ctypes.c_longlong), ("gave_up", ctypes.c_int),
                ("last_dt", ctypes.c_double), ("last_err", ctypes.c_double)]


CALLBACK = ctypes.CFUNCTYPE(None, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                            ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)

# int filter(const int* index, int n_dim, int depth, void* user)
FILTER = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                          ctypes.c_int, ctypes.c_void_p)

_lib = None


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libheom_amd: %s is missing; run libheom_amd.build.build() "
                          "(or __graft_entry__.build()) first" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, dp, llp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_longlong
    P = ctypes.POINTER
    sig = {
        "heom_last_error": (ctypes.c_char_p, []),
        "heom_version": (ip, [P(ip), P(ip)]),
        "heom_default_options": (ip, [P(heom_options)]),
        "heom_create": (ip, [ip, ip, ip, ip, P(heom_options), P(vp)]),
        "heom_destroy": (ip, [vp]),
        "heom_set_H": (ip, [vp, vp]),
        "heom_alloc_noises": (ip, [vp, ip]),
        "heom_set_noise": (ip, [vp, ip, vp, ip, vp, vp, vp, vp, vp, dp, dp]),
        "heom_set_hrchy_filter": (ip, [vp, FILTER, vp]),
        "heom_set_param": (ip, [vp]),
        "heom_get_n_hrchy": (ip, [vp, P(ip)]),
        "heom_get_n_dim": (ip, [vp, P(ip)]),
        "heom_main_size": (ip, [vp, P(llp)]),
        "heom_get_hrchy": (ip, [vp, vp, vp, vp]),
        "heom_redfield_lambda": (ip, [vp, ip, vp, vp]),
        "heom_calc_diff": (ip, [vp, vp, vp, dp, dp, dp, dp, vp]),
        "heom_calc_diff_host": (ip, [vp, vp, vp, dp, dp, dp, dp]),
        "heom_batch_create": (ip, [vp, ip, P(vp)]),
        "heom_batch_layout": (ip, [vp, vp, vp, vp]),
        "heom_batch_tuning": (ip, [vp, vp, vp, vp, ip, vp]),
        "heom_batch_set_tuning": (ip, [vp, ip, ip, ip]),
        "heom_solver_create": (ip, [vp, ip, P(vp)]),
        "heom_solver_destroy": (ip, [vp]),
        "heom_solve": (ip, [vp, vp, vp, ip, CALLBACK, vp, dp, dp, dp, ip]),
        "heom_solve_dev": (ip, [vp, vp, vp, ip, dp, dp, dp, vp]),
        "heom_solve_dev_noalloc": (ip, [vp, vp, vp, ip, dp, dp, dp, vp]),
        "heom_solver_stats": (ip, [vp, P(heom_stats)]),
        "heom_solver_reset": (ip, [vp]),
        "heom_solver_set_graphs": (ip, [vp, ip]),
        "heom_solver_time_stages": (ip, [vp, ip, vp, P(ip)]),
        "heom_solver_stage_bytes": (ip, [vp, vp, P(ip)]),
        "heom_dev_alloc": (ip, [P(vp), llp]),
        "heom_dev_free": (ip, [vp]),
        "heom_memcpy_h2d": (ip, [vp, vp, llp]),
        "heom_memcpy_d2h": (ip, [vp, vp, llp]),
        "heom_device_sync": (ip, []),
        "heom_stream_copy": (ip, [llp, ip, ip, P(dp)]),
        "heom_mlp_create": (ip, [ip, ip, ip, ctypes.c_ulonglong, ip, P(vp)]),
        "heom_mlp_destroy": (ip, [vp]),
        "heom_mlp_n_params": (ip, [vp, P(llp)]),
        "heom_mlp_get_params": (ip, [vp, vp]),
        "heom_mlp_set_params": (ip, [vp, vp]),
        "heom_mlp_set_scaling": (ip, [vp, vp, vp, vp, vp]),
        "heom_mlp_get_scaling": (ip, [vp, vp, vp, vp, vp]),
        "heom_mlp_fit_scaling": (ip, [vp, vp, vp, llp]),
        "heom_mlp_forward": (ip, [vp, vp, vp, ip]),
        "heom_mlp_loss_grad": (ip, [vp, vp, vp, ip, P(dp), vp]),
        "heom_mlp_adam_step": (ip, [vp, dp, dp, dp, dp]),
        "heom_mlp_train": (ip, [vp, vp, vp, llp, ip, ip, dp, ctypes.c_ulonglong, vp]),
        "heom_set_closure": (ip, [vp, vp, vp, ip, ip]),
        "heom_closure_dims": (ip, [vp, ip, ip, P(ip), P(ip), P(ip)]),
        "heom_closure_rows": (ip, [vp, vp, vp, vp, ip, ip, vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != HEOM_OK:
        msg = load().heom_last_error()
        raise HeomError(rc, msg.decode() if msg else "")
    return rc
