"""Loader for libsmectic.so (the HIP product library).

There is no fallback: if the library is missing or fails to load, every entry
point raises.  Build it with ``python -m smectic3d_amd.build``.
"""
import ctypes as C
import os

from ._abi import SmxConfig, SmxSchedule, SmxSurfOpts

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libsmectic.so")

# Every entry point declared in include/smectic.h.
EXPORTS = [
    "smx_config_default", "smx_create", "smx_destroy", "smx_last_error", "smx_get_config",
    "smx_set_psi", "smx_get_psi", "smx_step", "smx_l1", "smx_energy", "smx_dominant_wavenumber",
    "smx_dtpsi", "smx_surftrack", "smx_sync", "smx_step_count", "smx_profile_step",
    "smx_calibration_copy", "smx_adjoint_grad", "smx_batch_create", "smx_batch_destroy",
    "smx_batch_handle", "smx_batch_run", "smx_batch_graph_nodes", "smx_get_schedule", "smx_set_schedule",
    "smx_autotune", "smx_batch_autotune", "smx_set_adjoint_budget", "smx_adjoint_info",
    "smx_save_checkpoint", "smx_load_checkpoint", "smx_psi_avg", "smx_eckhaus_scan",
    "smx_set_l1_tracking", "smx_set_finite_check", "smx_batch_adjoint_grad", "smx_batch_adjoint_nodes",
    "smx_step_noalloc",
]

_lib = None
_dp = C.POINTER(C.c_double)


class SmxError(RuntimeError):
    def __init__(self, code, msg):
        from ._abi import ERR_NAMES
        super().__init__(f"{ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m smectic3d_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.smx_config_default.argtypes = [C.POINTER(SmxConfig), C.c_int]
    L.smx_config_default.restype = None
    L.smx_create.argtypes = [C.POINTER(SmxConfig), C.POINTER(vp)]
    L.smx_destroy.argtypes = [vp]
    L.smx_destroy.restype = None
    L.smx_last_error.argtypes = []
    L.smx_last_error.restype = C.c_char_p
    L.smx_get_config.argtypes = [vp, C.POINTER(SmxConfig)]
    for n in ("smx_set_psi", "smx_get_psi", "smx_dtpsi"):
        getattr(L, n).argtypes = [vp, _dp, C.c_int64]
    L.smx_step.argtypes = [vp, C.c_int64]
    L.smx_step_noalloc.argtypes = [vp, C.c_int64]
    for n in ("smx_l1", "smx_energy", "smx_dominant_wavenumber"):
        getattr(L, n).argtypes = [vp, _dp]
    L.smx_surftrack.argtypes = [vp, C.POINTER(SmxSurfOpts), _dp, _dp, _dp, _dp, C.POINTER(C.c_int64)]
    L.smx_psi_avg.argtypes = [vp, C.c_int, _dp, C.c_int64]
    L.smx_eckhaus_scan.argtypes = [vp, C.c_int, _dp, _dp, _dp]
    L.smx_sync.argtypes = [vp]
    L.smx_step_count.argtypes = [vp, C.POINTER(C.c_int64)]
    L.smx_profile_step.argtypes = [vp, C.c_int64, C.c_int, C.POINTER(C.c_int), _dp, _dp, C.c_char_p,
                                   C.c_int]
    L.smx_calibration_copy.argtypes = [vp, C.c_int64, C.c_int, _dp]
    L.smx_adjoint_grad.argtypes = [vp, _dp, C.c_int64, C.c_int64, C.c_int64, C.c_double, _dp, _dp, _dp, _dp]
    L.smx_batch_create.argtypes = [C.c_int, C.POINTER(SmxConfig), C.POINTER(vp)]
    L.smx_batch_destroy.argtypes = [vp]
    L.smx_batch_destroy.restype = None
    L.smx_batch_handle.argtypes = [vp, C.c_int, C.POINTER(vp)]
    L.smx_batch_run.argtypes = [vp, C.c_int64, C.c_int64]
    L.smx_batch_graph_nodes.argtypes = [vp, C.POINTER(C.c_int64)]
    L.smx_get_schedule.argtypes = [vp, C.POINTER(SmxSchedule)]
    L.smx_set_schedule.argtypes = [vp, C.POINTER(SmxSchedule)]
    L.smx_autotune.argtypes = [vp, C.c_int, C.POINTER(SmxSchedule)]
    L.smx_batch_autotune.argtypes = [vp, C.c_int]
    L.smx_set_adjoint_budget.argtypes = [vp, C.c_int64]
    L.smx_save_checkpoint.argtypes = [vp, C.c_char_p]
    L.smx_load_checkpoint.argtypes = [vp, C.c_char_p]
    L.smx_adjoint_info.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.smx_set_l1_tracking.argtypes = [vp, C.c_int]
    L.smx_set_finite_check.argtypes = [vp, C.c_int64]
    L.smx_batch_adjoint_grad.argtypes = [vp, C.POINTER(_dp), C.c_int64, C.c_int64, C.c_int64, _dp, _dp,
                                         C.POINTER(_dp), _dp, _dp]
    L.smx_batch_adjoint_nodes.argtypes = [vp, C.POINTER(C.c_int64)]
    for n in EXPORTS:
        f = getattr(L, n)
        if n not in ("smx_config_default", "smx_destroy", "smx_last_error", "smx_batch_destroy"):
            f.restype = C.c_int
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise SmxError(rc, lib().smx_last_error().decode())
    return rc
