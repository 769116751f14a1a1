"""ctypes binding of libibpm (include/ibpm.h).

This is the same binding a Julia ``ccall`` shim makes: plain pointers and sizes. The
library is loaded from the package's ``lib/`` folder; nothing falls back to a CPU path
when it is missing.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "lib", "libibpm.so")

IBPM_OK = 0
IBPM_E_ARG = 1
IBPM_E_OUTSIDE_REGION = 2
IBPM_E_HIP = 3
IBPM_E_SINGULAR = 4
IBPM_E_NO_DEVICE = 5

IBPM_FP64 = 0
IBPM_MIXED = 1

FIELDS = {"q": 0, "q0": 1, "Gamma": 2, "psi": 3, "nonlin1": 4, "nonlin2": 5, "xb": 6,
          "ub": 7, "Ftilde": 8, "fb": 9, "total_force": 10, "ds": 11, "freestream": 12}

# every symbol include/ibpm.h declares
EXPORTS = ("ibpm_default_options", "ibpm_create", "ibpm_advance", "ibpm_read", "ibpm_write",
           "ibpm_field_size", "ibpm_get_index", "ibpm_set_index", "ibpm_get_time",
           "ibpm_synchronize", "ibpm_last_error", "ibpm_destroy", "ibpm_time_kernel", "ibpm_apply",
           "ibpm_set_warm_start", "ibpm_net_default_config", "ibpm_net_create",
           "ibpm_net_destroy", "ibpm_net_dataset_size", "ibpm_net_clear_dataset",
           "ibpm_net_train", "ibpm_net_param_count", "ibpm_net_get_params",
           "ibpm_net_set_params", "ibpm_net_forward", "ibpm_net_gradients", "ibpm_net_step",
           "ibpm_batch_create", "ibpm_batch_advance", "ibpm_batch_destroy", "ibpm_dst_tiles", "ibpm_set_body_motion",
           "ibpm_box_update", "ibpm_apply_noalloc")

IBPM_WS_PREVIOUS = 0
IBPM_WS_EXTRAPOLATE = 1
IBPM_WS_LEARNED = 2
NET_FEATURES = 10
NET_OUTPUTS = 2

IBPM_OP_POISSON = 0
IBPM_OP_HELMHOLTZ = 1
IBPM_OP_VORT2FLUX = 2
NOALLOC_SCRATCH_BYTES = 32 << 20  # IBPM_NOALLOC_SCRATCH_BYTES

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Problem(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nlevel", C.c_int),
                ("x0", C.c_double), ("y0", C.c_double), ("h", C.c_double),
                ("Re", C.c_double), ("dt", C.c_double), ("t0", C.c_double),
                ("npanel", C.c_int), ("xb", _dp), ("ds", _dp), ("nbody", C.c_int),
                ("body_start", _ip), ("body_count", _ip), ("body_moving", _ip),
                ("grid_moving", C.c_int)]


class Options(C.Structure):
    _fields_ = [("precision", C.c_int), ("refine_sweeps", C.c_int), ("cg_maxiter", C.c_int),
                ("cg_reltol", C.c_double), ("cg_abstol", C.c_double), ("cg_tol_mode", C.c_int),
                ("device", C.c_int), ("dst_tiles", C.c_int), ("cg_precond", C.c_int),
                ("cg_precond_refresh", C.c_int), ("helmholtz", C.c_int),
                ("box_update", C.c_int)]


class StepStats(C.Structure):
    _fields_ = [("cg_iters", C.c_int), ("cg_converged", C.c_int), ("cg_residual", C.c_double),
                ("force", C.c_double * 2), ("b_applies", C.c_int), ("ws_used", C.c_int),
                ("ws_base_residual", C.c_double), ("ws_residual", C.c_double),
                ("pc_assembled", C.c_int)]


class NetConfig(C.Structure):
    _fields_ = [("channels", C.c_int), ("layers", C.c_int), ("kernel", C.c_int),
                ("max_points", C.c_int), ("capacity", C.c_int), ("max_batch", C.c_int),
                ("seed", C.c_ulonglong), ("device", C.c_int)]


class IbpmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


_lib = None


def load():
    """Load libibpm.so; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libibpm.so not built: {LIB_PATH} is missing "
                          "(run __graft_entry__.build() or make -C immersedbodies.jl_amd)")
    L = C.CDLL(LIB_PATH)
    L.ibpm_default_options.argtypes = [C.POINTER(Options)]
    L.ibpm_default_options.restype = None
    L.ibpm_create.argtypes = [C.POINTER(Problem), C.POINTER(Options), C.POINTER(C.c_void_p)]
    L.ibpm_create.restype = C.c_int
    L.ibpm_advance.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.POINTER(StepStats)]
    L.ibpm_advance.restype = C.c_int
    L.ibpm_read.argtypes = [C.c_void_p, C.c_int, _dp, C.c_size_t]
    L.ibpm_read.restype = C.c_int
    L.ibpm_write.argtypes = [C.c_void_p, C.c_int, _dp, C.c_size_t]
    L.ibpm_write.restype = C.c_int
    L.ibpm_field_size.argtypes = [C.c_void_p, C.c_int]
    L.ibpm_field_size.restype = C.c_size_t
    L.ibpm_get_index.argtypes = [C.c_void_p]
    L.ibpm_get_index.restype = C.c_int
    L.ibpm_set_index.argtypes = [C.c_void_p, C.c_int]
    L.ibpm_set_index.restype = C.c_int
    L.ibpm_get_time.argtypes = [C.c_void_p]
    L.ibpm_get_time.restype = C.c_double
    L.ibpm_synchronize.argtypes = [C.c_void_p]
    L.ibpm_synchronize.restype = C.c_int
    L.ibpm_last_error.argtypes = []
    L.ibpm_last_error.restype = C.c_char_p
    L.ibpm_destroy.argtypes = [C.c_void_p]
    L.ibpm_destroy.restype = None
    L.ibpm_time_kernel.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp]
    L.ibpm_time_kernel.restype = C.c_int
    L.ibpm_apply.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, C.c_size_t]
    L.ibpm_apply.restype = C.c_int
    # device pointers
    L.ibpm_apply_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                     C.c_size_t]
    L.ibpm_apply_noalloc.restype = C.c_int
    _fp = C.POINTER(C.c_float)
    L.ibpm_set_warm_start.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int]
    L.ibpm_set_warm_start.restype = C.c_int
    L.ibpm_net_default_config.argtypes = [C.POINTER(NetConfig)]
    L.ibpm_net_default_config.restype = None
    L.ibpm_net_create.argtypes = [C.POINTER(NetConfig), C.POINTER(C.c_void_p)]
    L.ibpm_net_create.restype = C.c_int
    L.ibpm_net_destroy.argtypes = [C.c_void_p]
    L.ibpm_net_destroy.restype = None
    L.ibpm_net_dataset_size.argtypes = [C.c_void_p]
    L.ibpm_net_dataset_size.restype = C.c_int
    L.ibpm_net_clear_dataset.argtypes = [C.c_void_p]
    L.ibpm_net_clear_dataset.restype = C.c_int
    L.ibpm_net_train.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, _dp]
    L.ibpm_net_train.restype = C.c_int
    L.ibpm_net_param_count.argtypes = [C.c_void_p]
    L.ibpm_net_param_count.restype = C.c_size_t
    L.ibpm_net_get_params.argtypes = [C.c_void_p, _fp, C.c_size_t]
    L.ibpm_net_get_params.restype = C.c_int
    L.ibpm_net_set_params.argtypes = [C.c_void_p, _fp, C.c_size_t]
    L.ibpm_net_set_params.restype = C.c_int
    L.ibpm_net_forward.argtypes = [C.c_void_p, _fp, _ip, C.c_int, _fp]
    L.ibpm_net_forward.restype = C.c_int
    L.ibpm_net_gradients.argtypes = [C.c_void_p, _fp, _ip, _fp, C.c_int, _fp, C.c_size_t, _dp]
    L.ibpm_net_gradients.restype = C.c_int
    L.ibpm_net_step.argtypes = [C.c_void_p, _fp, _ip, _fp, C.c_int, C.c_double, _dp]
    L.ibpm_net_step.restype = C.c_int
    L.ibpm_batch_create.argtypes = [C.c_int, C.POINTER(C.c_void_p), _dp, C.POINTER(C.c_void_p)]
    L.ibpm_batch_create.restype = C.c_int
    L.ibpm_batch_advance.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, C.POINTER(StepStats)]
    L.ibpm_batch_advance.restype = C.c_int
    L.ibpm_batch_destroy.argtypes = [C.c_void_p]
    L.ibpm_batch_destroy.restype = None
    L.ibpm_dst_tiles.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _ip, _ip, _ip, _ip]
    L.ibpm_dst_tiles.restype = C.c_int
    L.ibpm_set_body_motion.argtypes = [C.c_void_p, C.c_int, C.c_double, _dp, _dp, C.c_double,
                                       C.c_double]
    L.ibpm_set_body_motion.restype = C.c_int
    L.ibpm_box_update.argtypes = [C.c_void_p, C.c_int, _ip]
    L.ibpm_box_update.restype = C.c_int
    _lib = L
    return L


def check(rc: int):
    if rc != IBPM_OK:
        msg = load().ibpm_last_error()
        raise IbpmError(rc, msg.decode() if msg else f"libibpm error {rc}")
