"""ctypes binding of libplbm.so (include/plbm.h).

The library is loaded from this package directory; a missing library is an error
(there is no CPU fallback for any engine call).
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIBPATH = PKG / "libplbm.so"

# status codes and enums, include/plbm.h
OK = 0
ERR_ARG, ERR_HIP, ERR_STATE, ERR_NOMEM, ERR_NUMERIC, ERR_TIMEOUT, ERR_HOST = 1, 2, 3, 4, 5, 6, 7
BGK, TRT, RR, BGK_SPLIT = 0, 1, 2, 3
F64, F64S, F32S, F16S = 0, 1, 2, 3
RED_MASS, RED_KE, RED_ERR2, RED_REF2, RED_UMAX, RED_NONFINITE, RED_U2 = range(7)
NRED = 8

MODEL_IDS = {"bgk": BGK, "trt": TRT, "rr": RR, "bgk_split": BGK_SPLIT}
STORAGE_IDS = {"f64": F64, "f64s": F64S, "f32s": F32S, "f16s": F16S}
STORAGE_BYTES = {F64: 8, F64S: 8, F32S: 4, F16S: 2}


class PlbmError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"plbm status {status}: {msg}")
        self.status = status


class LatticeInfo(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("ldd", ctypes.c_int), ("nf", ctypes.c_int),
        ("storage", ctypes.c_int), ("device", ctypes.c_int), ("omega", ctypes.c_double),
        ("trt_magic", ctypes.c_double), ("lambda_odd", ctypes.c_double),
        ("f16_scale", ctypes.c_double), ("bytes_per_buffer", ctypes.c_size_t),
    ]


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_i, _d, _sz, _u64 = ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_uint64

# name: (restype, argtypes)
SIGNATURES = {
    "plbm_create": (_i, [_i, _i, _i, _i, ctypes.POINTER(_vp)]),
    "plbm_destroy": (_i, [_vp]),
    "plbm_get_info": (_i, [_vp, ctypes.POINTER(LatticeInfo)]),
    "plbm_set_stream": (_i, [_vp, _vp]),
    "plbm_buffer_ptr": (_i, [_vp, _i, ctypes.POINTER(_vp)]),
    "plbm_set_params": (_i, [_vp, _d, _d]),
    "plbm_set_f16_scale": (_i, [_vp, _d]),
    "plbm_upload_f": (_i, [_vp, _i, _dp, _i]),
    "plbm_download_f": (_i, [_vp, _i, _dp, _i]),
    "plbm_equilibrium_init": (_i, [_vp, _i, _dp, _dp, _dp]),
    "plbm_equilibrium_init_dev": (_i, [_vp, _i, _vp, _vp, _vp]),
    "plbm_step": (_i, [_vp, _i, _i, _i]),
    "plbm_run": (_i, [_vp, _i, _i, _ip, _ip]),
    "plbm_stream": (_i, [_vp, _i, _i]),
    "plbm_collide": (_i, [_vp, _i, _i]),
    "plbm_copy": (_i, [_vp, _i, _i]),
    "plbm_macros": (_i, [_vp, _i, _dp, _dp, _dp]),
    "plbm_macros_dev": (_i, [_vp, _i, _vp, _vp, _vp]),
    "plbm_reduce": (_i, [_vp, _i, _dp, _dp, _dp]),
    "plbm_reduce_dev": (_i, [_vp, _i, _vp, _vp, _dp]),
    "plbm_reduce_static": (_i, [_vp, _i, _vp, _vp, _vp]),
    "plbm_sync": (_i, [_vp]),
    "plbm_last_error": (ctypes.c_char_p, []),
    "plbm_field_taylor_green": (_i, [_i, _i, _d, _d, _d, _d, _d, _dp, _dp, _dp]),
    "plbm_field_shear_layer": (_i, [_i, _i, _d, _d, _d, _d, _dp, _dp, _dp]),
    "plbm_field_add_noise": (_i, [_i, _u64, _d, _dp, _dp]),
    "plbm_noise": (_d, [_u64, _u64]),
    "plbm_lambda_d": (_d, [_d, _d]),
    "plbm_magic_number": (_d, [_d, _d]),
    "plbm_copy_bandwidth": (_i, [_sz, _i, _dp]),
    "plbm_set_launch": (_i, [_vp, _i, _i, _i]),
    "plbm_get_launch": (_i, [_vp, _i, _ip, _ip]),
    "plbm_autotune": (_i, [_vp, _i, _i, _i, _dp]),
    "plbm_batch_create": (_i, [_i, _ip, _ip, _i, _i, ctypes.POINTER(_vp)]),
    "plbm_batch_destroy": (_i, [_vp]),
    "plbm_batch_info": (_i, [_vp, _ip, ctypes.POINTER(ctypes.c_longlong),
                             ctypes.POINTER(_sz)]),
    "plbm_batch_set_params": (_i, [_vp, _dp, _dp]),
    "plbm_batch_equilibrium_init": (_i, [_vp, _i, _dp, _dp, _dp]),
    "plbm_batch_upload_f": (_i, [_vp, _i, _dp, _i]),
    "plbm_batch_download_f": (_i, [_vp, _i, _dp, _i]),
    "plbm_batch_macros": (_i, [_vp, _i, _dp, _dp, _dp]),
    "plbm_batch_reduce": (_i, [_vp, _dp]),
    "plbm_batch_run": (_i, [_vp, _i]),
    "plbm_batch_run_direct": (_i, [_vp, _i]),
    "plbm_batch_sync": (_i, [_vp]),
    "plbm_batch_set_launch": (_i, [_vp, _i, _i, _i, _i]),
    "plbm_batch_get_launch": (_i, [_vp, _ip, _ip, _ip, _ip]),
    "plbm_batch_set_steps_per_pass": (_i, [_vp, _i]),
    "plbm_batch_get_steps_per_pass": (_i, [_vp, _ip]),
    "plbm_batch_set_pass_columns": (_i, [_vp, _i]),
    "plbm_batch_get_pass_columns": (_i, [_vp, _ip]),
    "plbm_batch_autotune": (_i, [_vp, _i, _dp]),
    "plbm_adjoint_create": (_i, [_i, _i, _i, _i, _i, ctypes.POINTER(_vp)]),
    "plbm_adjoint_destroy": (_i, [_vp]),
    "plbm_adjoint_set_stream": (_i, [_vp, _vp]),
    "plbm_adjoint_forward_dev": (_i, [_vp, _vp, _vp, _vp, _d, _vp, _vp, _dp]),
    "plbm_adjoint_backward_dev": (_i, [_vp, _vp, _vp, _vp, _dp]),
    "plbm_field_band_limited": (_i, [_i, _i, _i, _d, _u64, _dp, _dp]),
    "plbm_stream_op": (_i, [_vp, _i, _i, _i, _d]),
    "plbm_step_op": (_i, [_vp, _i, _i, _i, _i, _d]),
    "plbm_run_op": (_i, [_vp, _i, _i, _i, _ip, _ip, _d]),
    "plbm_set_op_steps_per_pass": (_i, [_vp, _i]),
    "plbm_get_op_steps_per_pass": (_i, [_vp, _i, _i, _ip]),
    "plbm_set_steps_per_launch": (_i, [_vp, _i, _i]),
    "plbm_pass": (_i, [_vp, _i, _i, _i, _i]),
    "plbm_get_steps_per_launch": (_i, [_vp, _i, _ip]),
    "plbm_set_pass_segments": (_i, [_vp, _i]),
    "plbm_get_pass_segments": (_i, [_vp, _ip]),
    "plbm_set_pass_lag_buffer": (_i, [_vp, _i]),
    "plbm_set_run_timing": (_i, [_vp, _i]),
    "plbm_run_launch_times": (_i, [_vp, _i, _ip, ctypes.POINTER(ctypes.c_float), _ip]),
    "plbm_dugks_collide": (_i, [_vp, _i, _i, _d, _d]),
    "plbm_dugks_stream": (_i, [_vp, _i, _i, _d]),
    "plbm_dugks_step": (_i, [_vp, _i, _i, _i, _d, _d, _d]),
    "plbm_dugks_set_steps_per_pass": (_i, [_vp, _i]),
    "plbm_dugks_get_steps_per_pass": (_i, [_vp, _ip]),
    "plbm_dugks_stream_faces": (_i, [_vp, _i, _i, _d, _d]),
    "plbm_dugks_step_faces": (_i, [_vp, _i, _i, _i, _d, _d, _d, _d]),
    "plbm_field_barotropic_vortex": (_i, [_i, _i, _d, _d, _d, _d, _d, _d, _d, _dp, _dp, _dp]),
}
STREAM_OPS = {"exact": 0, "fvm_bardow": 1, "fdm_bardow": 2, "fdm_sofonea": 3,
              "fdm_bardow_wls": 4, "fdm_bardow_wls_gauss_v1": 5, "fdm_bardow_wls_gauss_v2": 6,
              "fdm_bardow_iso": 7}

_LIB = None


def lib() -> ctypes.CDLL:
    """Load libplbm.so once; raise if it has not been built."""
    global _LIB
    if _LIB is None:
        if not LIBPATH.exists():
            raise ImportError(
                f"{LIBPATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                f"g.build()'` (there is no CPU fallback)")
        L = ctypes.CDLL(str(LIBPATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(status: int) -> None:
    if status != OK:
        msg = lib().plbm_last_error()
        raise PlbmError(status, msg.decode() if msg else "")


def dptr(a: np.ndarray):
    if a.dtype != np.float64:
        raise TypeError("expected float64 array")
    if not (a.flags.f_contiguous or a.ndim <= 1):
        raise ValueError("expected a Fortran-contiguous array")
    return a.ctypes.data_as(_dp)


def exported_symbols() -> list[str]:
    return sorted(SIGNATURES)
