"""ctypes binding of libmhd3d.so (the C ABI of include/mhd3d.h).

The library is the product: it must exist (built by ``build.py`` /
``__graft_entry__.build()``); there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "libmhd3d.so")

OK, EINVAL, EUNSUPPORTED, EHIP, ENOMEM, ENONFINITE, EHALTED, EEXIT = range(8)
REFERENCE_DT_MIN = float(np.float32(1.0e-8))   # the reference driver's DT guard (DT < 1.0e-8f)
FP64, MIXED = 0, 1
FIELDS = ("RU", "RV", "RW", "RO", "TT", "BX", "BY", "BZ")
FLUXES = ("FU", "FV", "FW", "FR", "FT", "WW1", "WW2", "WW3")

# Every entry point the header declares: name -> (restype, argtypes)
_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_dpp = C.POINTER(_dp)
SIGNATURES = {
    "mhd3d_create": (C.c_int, [_vp, C.POINTER(_vp)]),
    "mhd3d_destroy": (C.c_int, [_vp]),
    "mhd3d_set_state": (C.c_int, [_vp, _dpp]),
    "mhd3d_get_state": (C.c_int, [_vp, _dpp]),
    "mhd3d_step": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    "mhd3d_step_noalloc": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    "mhd3d_fluxes": (C.c_int, [_vp, _dpp]),
    "mhd3d_stage": (C.c_int, [_vp, C.c_int, C.c_double]),
    "mhd3d_bcon": (C.c_int, [_vp]),
    "mhd3d_cfl": (C.c_int, [_vp, _vp]),
    "mhd3d_diagnostics": (C.c_int, [_vp, _dp]),
    "mhd3d_clock": (C.c_int, [_vp, C.POINTER(C.c_int64), _dp, _dp]),
    "mhd3d_relax_state": (C.c_int, [_vp, _dp, C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "mhd3d_autotune": (C.c_int, [_vp, C.c_int, _dp, C.POINTER(C.c_int)]),
    "mhd3d_get_variant": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "mhd3d_timing_enable": (C.c_int, [_vp, C.c_int]),
    "mhd3d_timing_read": (C.c_int, [_vp, _vp]),
    "mhd3d_synchronize": (C.c_int, [_vp]),
    "mhd3d_last_error": (C.c_char_p, [_vp]),
    "mhd3d_device_bytes": (C.c_int64, [_vp]),
    "mhd3d_set_dt_guard": (C.c_int, [_vp, C.c_double]),
    "mhd3d_geometry": (C.c_int, [_vp, C.POINTER(C.c_int32)]),
    "mhd3d_static_state": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, C.c_int, _dpp]),
    "mhd3d_static_state_flags": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, C.c_int,
                                           C.c_int, C.c_int, _dpp, _dp, _dp, _dp]),
    "mhd3d_dump_data": (C.c_int, [_vp, C.c_char_p, C.c_int]),
    "mhd3d_load_data": (C.c_int, [_vp, C.c_char_p, C.c_int]),
    "mhd3d_dump_param": (C.c_int, [C.c_char_p, C.c_int, _dp]),
    "mhd3d_load_param": (C.c_int, [C.c_char_p, C.c_int, _dp]),
    "mhd3d_calibration_copy": (C.c_int, [C.c_int64, C.c_int, _dp]),
    "mhd3d_stream_ceiling": (C.c_int, [C.c_int64, C.c_int, C.c_int, C.c_int, _dp]),
    "mhd3d_ensemble_set_dt_guard": (C.c_int, [_vp, C.c_double]),
    "mhd3d_ensemble_set_dt_guard_member": (C.c_int, [_vp, C.c_int, C.c_double]),
    "mhd3d_ensemble_halted": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int)]),
    "mhd3d_init_stratified": (C.c_int, [_vp, _dp, _dp, C.c_uint64, C.c_double, C.c_double, C.c_double]),
    "mhd3d_ensemble_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.POINTER(_vp)]),
    "mhd3d_ensemble_destroy": (C.c_int, [_vp]),
    "mhd3d_ensemble_last_error": (C.c_char_p, [_vp]),
    "mhd3d_ensemble_size": (C.c_int, [_vp]),
    "mhd3d_ensemble_cells": (C.c_int64, [_vp]),
    "mhd3d_ensemble_set_state": (C.c_int, [_vp, C.c_int, _dpp]),
    "mhd3d_ensemble_get_state": (C.c_int, [_vp, C.c_int, _dpp]),
    "mhd3d_ensemble_step": (C.c_int, [_vp, C.c_int]),
    "mhd3d_ensemble_clock": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int64), _dp, _dp, _dp]),
    "mhd3d_ensemble_graph": (C.c_int, [_vp, C.c_int]),
    "mhd3d_ensemble_set_variant": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "mhd3d_ensemble_get_variant": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "mhd3d_ensemble_synchronize": (C.c_int, [_vp]),
    "mhd3d_ensemble_autotune": (C.c_int, [_vp, C.c_int, _dp, C.POINTER(C.c_int)]),
    "mhd3d_server_create": (C.c_int, [C.c_int, C.c_int64, C.c_int, C.POINTER(_vp)]),
    "mhd3d_server_destroy": (C.c_int, [_vp]),
    "mhd3d_server_last_error": (C.c_char_p, [_vp]),
    "mhd3d_server_now": (C.c_double, [_vp]),
    "mhd3d_server_bytes_in_use": (C.c_int64, [_vp]),
    "mhd3d_server_cell_steps": (C.c_int64, [_vp]),
    "mhd3d_server_submit": (C.c_int, [_vp, _vp, _dp, _dp, C.c_uint64, C.c_double, C.c_double,
                                      C.c_double, C.c_int, C.POINTER(C.c_int64)]),
    "mhd3d_server_tick": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "mhd3d_server_poll": (C.c_int, [_vp, C.c_int64, _vp]),
    "mhd3d_server_result": (C.c_int, [_vp, C.c_int64, _dpp]),
}

REQ_EXITED, REQ_NONFINITE, REQ_FAILED, REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_RELEASED = -3, -2, -1, 0, 1, 2, 3


class RequestInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("steps_done", C.c_int32), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_finish", C.c_double), ("dt_last", C.c_double)]


class CflInfo(C.Structure):
    _fields_ = [("wmin", C.c_double * 4), ("umach", C.c_double), ("dt", C.c_double)]


class Timing(C.Structure):
    _fields_ = [("stage_ms_total", C.c_double), ("stage_launches", C.c_int64),
                ("step_ms_total", C.c_double), ("steps", C.c_int64)]


_lib = None


class MhdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mhd3d error {code}: {msg}")
        self.code = code


def lib():
    """Load libmhd3d.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                               "or rast-3dmhd-cpp_amd/build.py first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _fieldptrs(arrays):
    arr = (_dp * len(arrays))()
    for k, a in enumerate(arrays):
        arr[k] = a.ctypes.data_as(_dp)
    return arr
