"""Loader/builder for the oracle's C helpers (oracle/spread.c).  Test
infrastructure only."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SRC = os.path.join(_HERE, "spread.c")
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile spread.c with gcc/OpenMP into oracle/liboracle.so."""
    if force or not os.path.exists(_LIB) or os.path.getmtime(_LIB) < os.path.getmtime(_SRC):
        tmp = _LIB + ".tmp"
        subprocess.check_call(["gcc", "-O3", "-march=x86-64-v2", "-fopenmp", "-fPIC", "-shared",
                               _SRC, "-o", tmp, "-lm"])
        os.replace(tmp, _LIB)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        p = ctypes.c_void_p
        i64 = ctypes.c_int64
        L.es_spread_f64.argtypes = [ctypes.c_int, i64, p, ctypes.c_int, p, p, ctypes.c_int,
                                    ctypes.c_double, p]
        L.es_interp_f64.argtypes = [ctypes.c_int, i64, p, ctypes.c_int, p, p, ctypes.c_int,
                                    ctypes.c_double, p]
        L.spread_f32_1ch.argtypes = [ctypes.c_int, i64, p, p, p, ctypes.c_int, ctypes.c_float, p,
                                     p, ctypes.c_int]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def es_spread_f64(s, vals, n, w, beta):
    """s: (M, d) float64 grid coords; vals: (nch, M) complex128; n: dims with
    n[0] fastest.  Returns (nch, prod(n)) complex128."""
    s = np.ascontiguousarray(s, dtype=np.float64)
    vals = np.ascontiguousarray(vals, dtype=np.complex128)
    M, d = s.shape
    nch = vals.shape[0]
    nn = np.ascontiguousarray(n, dtype=np.int64)
    out = np.empty((nch, int(np.prod(nn))), dtype=np.complex128)
    lib().es_spread_f64(d, M, _ptr(s), nch, _ptr(vals), _ptr(nn), int(w), float(beta), _ptr(out))
    return out


def es_interp_f64(s, grid, n, w, beta):
    s = np.ascontiguousarray(s, dtype=np.float64)
    grid = np.ascontiguousarray(grid, dtype=np.complex128)
    M, d = s.shape
    nch = grid.shape[0]
    nn = np.ascontiguousarray(n, dtype=np.int64)
    out = np.empty((nch, M), dtype=np.complex128)
    lib().es_interp_f64(d, M, _ptr(s), nch, _ptr(grid), _ptr(nn), int(w), float(beta), _ptr(out))
    return out


def spread_f32_1ch(s, vals, n, w, beta, scratch, nthreads):
    s = np.ascontiguousarray(s, dtype=np.float32)
    vals = np.ascontiguousarray(vals, dtype=np.complex64)
    M, d = s.shape
    nn = np.ascontiguousarray(n, dtype=np.int64)
    out = np.empty(int(np.prod(nn)), dtype=np.complex64)
    lib().spread_f32_1ch(d, M, _ptr(s), _ptr(vals), _ptr(nn), int(w), float(beta), _ptr(out),
                         _ptr(scratch), int(nthreads))
    return out
