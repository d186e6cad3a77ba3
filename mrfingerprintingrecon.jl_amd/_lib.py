"""ctypes binding of libmsr.so (the C ABI declared in include/msr.h) and the
mapping of its status codes to the exception types the reference throws."""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmsr.so")

MSR_OK = 0
MSR_ERR_ARGUMENT = 1
MSR_ERR_ASSERTION = 2
MSR_ERR_ERROR = 3
MSR_ERR_UNSUPPORTED = 4
MSR_ERR_NOMEM = 5
MSR_ERR_HIP = 6
MSR_ERR_IO = 7

DCF_MODES = {"none": 0, "radial_2D": 1, "radial_3D": 2}
DCF_ARRAY = 3
CG_MODES = {"device": 0, "host": 1, "graph": 2}
CG_PRECISIONS = {"f32": 0, "mixed": 1, "f64": 2, "f64c": 3}
TUNE_CONFIGS = 16


class ArgumentError(ValueError):
    """Julia ArgumentError (src/BackProjection.jl:183-202)."""


class ErrorException(RuntimeError):
    """Julia ErrorException (src/BackProjection.jl:179)."""


class HIPError(RuntimeError):
    pass


_EXC = {
    MSR_ERR_ARGUMENT: ArgumentError,
    MSR_ERR_ASSERTION: AssertionError,
    MSR_ERR_ERROR: ErrorException,
    MSR_ERR_UNSUPPORTED: NotImplementedError,
    MSR_ERR_NOMEM: MemoryError,
    MSR_ERR_HIP: HIPError,
    MSR_ERR_IO: OSError,
}

# exported symbols and their signatures (kept in step with include/msr.h)
_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_D = ctypes.c_double
SIGNATURES = {
    "msr_last_error": (ctypes.c_size_t, [ctypes.c_char_p, ctypes.c_size_t]),
    "msr_version": (_I, []),
    "msr_ctx_create": (_I, [_I, _P]),
    "msr_ctx_destroy": (_I, [_P]),
    "msr_traj_create": (_I, [_P, _I, _I64, _I64, _P, _P, _P, _P]),
    "msr_traj_info": (_I, [_P, _P, _P]),
    "msr_traj_destroy": (_I, [_P]),
    "msr_normalop_create_from_traj": (_I, [_P, _P, _I, _P, _I, _P, _I, _I, _P, _P, _P]),
    "msr_normalop_create": (_I, [_P, _I, _P, _I64, _I, _P, _P, _I64, _I, _P, _P, _P]),
    "msr_normalop_info": (_I, [_P, _P, _P, _P, _P]),
    "msr_normalop_export_kernel": (_I, [_P, _P, _P, _P]),
    "msr_normalop_apply": (_I, [_P, _P, _P, _D, _D, _D, _D, _P]),
    "msr_normalop_apply_percoil": (_I, [_P, _P, _P, _D, _D, _D, _D, _P]),
    "msr_normalop_destroy": (_I, [_P]),
    "msr_backprojection": (_I, [_P, _P, _I, _P, _P, _I, _I, _P, _I, _P, _I, _P, _P, _P]),
    "msr_backprojection_channelwise": (_I, [_P, _P, _I, _P, _P, _I, _I, _P, _I, _P, _I, _P, _P, _P]),
    "msr_forward": (_I, [_P, _P, _I, _P, _P, _I, _P, _I, _I, _P, _P, _P]),
    "msr_cg": (_I, [_P, _P, _P, _I, _D, _D, _I, _I, _I, _P, _P, _P]),
    "msr_cg_noalloc": (_I, [_P, _P, _P, _I, _D, _D, _I, _I, _P, _P, _P]),
    "msr_normalop_apply_f64": (_I, [_P, _P, _P, _D, _D, _D, _D, _P]),
    "msr_normalop_apply_f64c": (_I, [_P, _P, _P, _D, _D, _D, _D, _P]),
    "msr_normalop_tune": (_I, [_P, _I, _I, _P, _P, _P]),
    "msr_tune_save": (_I, [_P, ctypes.c_char_p]),
    "msr_tune_load": (_I, [_P, ctypes.c_char_p, _P]),
    "msr_normalop_create_shared": (_I, [_P, _P, _I, _P, _P, _P]),
    "msr_normalop_footprint": (_I, [_P, _I, _P, _P]),
    "msr_kernel_key": (_I, [_P, _P, _I, _P, _I, _P, _I, _P, _P]),
    "msr_kernel_save": (_I, [_P, ctypes.c_char_p, _P, _P]),
    "msr_kernel_load": (_I, [_P, ctypes.c_char_p, _P, _I, _P, _P, _P]),
    "msr_kernel_file_info": (_I, [ctypes.c_char_p, _I, _P, _P, _P, _P, _P, _P]),
    "msr_kernel_cache_create": (_I, [_P, ctypes.c_size_t, ctypes.c_char_p, _P]),
    "msr_kernel_cache_destroy": (_I, [_P]),
    "msr_kernel_cache_get": (_I, [_P, _P, _I, _P, _I, _P, _I, _I, _P, _P, _P, _P, _P]),
    "msr_kernel_cache_stats": (_I, [_P, _P, _P, _P]),
    "msr_traj_create_cartesian": (_I, [_P, _I, _I64, _I64, _P, _P, _P, _P]),
    "msr_fftnormalop_create_from_traj": (_I, [_P, _P, _I, _P, _I, _P, _I, _I, _P, _P, _P]),
    "msr_fftnormalop_create_from_mask": (_I, [_P, _I, _P, _I64, _P, _I, _P, _I, _I, _P, _P, _P]),
    "msr_fftnormalop_create": (_I, [_P, _I, _P, _I, _I, _P, _I, _P, _P, _P]),
    "msr_normalop_kind": (_I, [_P, _P]),
    "msr_calib_select": (_I, [_P, _I, _I64, _I64, _P, _P, _P, _P, _P, _P, _P]),
    "msr_reconstruct_coilwise": (_I, [_P, _P, _I, _P, _P, _I, _P, _I, _I, _I, _P, _P]),
    "msr_calib_kspace": (_I, [_P, _I, _P, _I, _P, _P, _P]),
    "msr_server_create": (_I, [_P, ctypes.c_size_t, _I, _I, _P]),
    "msr_server_destroy": (_I, [_P]),
    "msr_server_submit": (_I, [_P, _P, _P, _P, _I, _D, _D, _P]),
    "msr_server_step": (_I, [_P, _I, _P, _P, _P]),
    "msr_server_set_policy": (_I, [_P, _I]),
    "msr_server_poll": (_I, [_P, ctypes.c_int64, _P, _P, _P, _P]),
    "msr_server_forget": (_I, [_P, ctypes.c_int64]),
    "msr_batch_create": (_I, [_P, _I, _P, _P]),
    "msr_batch_destroy": (_I, [_P]),
    "msr_batch_apply": (_I, [_P, _P, _P, _D, _D, _D, _D, _P]),
    "msr_batch_cg": (_I, [_P, _P, _P, _I, _D, _D, _I, _I, _I, _P, _P, _P]),
    "msr_batch_cg_start": (_I, [_P, _P, _P, _I, _D, _D, _I, _I, _P]),
    "msr_batch_cg_iterate": (_I, [_P, _I, _P]),
    "msr_batch_cg_finish": (_I, [_P, _P]),
    "msr_batch_cg_status": (_I, [_P, _P, _P, _P, _P]),
}

_lib = None
_lock = threading.Lock()


def load(path: str = LIB_PATH):
    """Load libmsr.so.  Raises if it is missing: there is no fallback path."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(path):
                raise ImportError(
                    f"{path} is missing: build the HIP library first "
                    "(python mrfingerprintingrecon.jl_amd/build.py)")
            L = ctypes.CDLL(path)
            for name, (res, args) in SIGNATURES.items():
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
            _lib = L
    return _lib


def last_error() -> str:
    L = load()
    buf = ctypes.create_string_buffer(2048)
    L.msr_last_error(buf, 2048)
    return buf.value.decode(errors="replace")


def check(code: int):
    if code != MSR_OK:
        raise _EXC.get(code, RuntimeError)(last_error())


def call(name: str, *args):
    check(getattr(load(), name)(*args))
