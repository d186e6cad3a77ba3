"""MI355X-native Toeplitz subspace normal operator, subspace backprojection and
device CG: the host-side mirror of MRISubspaceRecon.jl's GPU API
(repo MagneticResonanceImaging/MRFingerprintingRecon.jl,
ext/MRISubspaceReconCUDAExt) over the C ABI of libmsr.so (include/msr.h).

Arrays are torch tensors on a HIP device in the row-major view of the
reference's Julia arrays (every shape reversed, identical memory):
  trj         Julia (d, Nsamp, Nt)       -> (Nt, Nsamp, d) float32
              4D    (d, Nr, Ncyc, Nt)     -> (Nt, Ncyc, Nr, d)
  sample_mask Julia (Nsamp, Nt)           -> (Nt, Nsamp) bool
  data        Julia (Nsamp, Nt, C)        -> (C, Nt, Nsamp) complex64
  U           Julia (Nt, K)               -> (K, Nt) complex64 or float32
  images      Julia (N1, N2[, N3], K)     -> (K[, N3], N2, N1) complex64
  cmaps       Vector of Julia (N1, N2...) -> list of (..., N2, N1) tensors or a
                                              stacked (C, ..., N2, N1) tensor
  packed Λ    Julia (P, Nkm)              -> (Nkm, P)
The operator's vectors are flat (Julia vec) tensors of length prod(N) * K.
There is no CPU path: every call runs HIP kernels and raises if libmsr.so is
missing.
"""
from __future__ import annotations

import ctypes
import math
import os

from . import _lib, synth  # noqa: F401
from ._lib import ArgumentError, ErrorException  # noqa: F401

__all__ = ["NFFTNormalOp", "calculate_backprojection", "forward", "cg", "cg_", "mul_",
           "LinearOperator", "plan_trajectory", "ArgumentError", "ErrorException", "synth",
           "kernel_key", "save_kernel", "load_kernel", "kernel_file_info", "KernelCache",
           "reconstruct_coilwise", "calibration_region", "calibration_kspace",
           "calculate_coil_maps", "default_calib_size", "FFTNormalOp", "save_tuning",
           "load_tuning"]

_ctxs = {}


def _np():
    import numpy
    return numpy


def _torch():
    import torch
    return torch


def _stream_ptr(device):
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _ctx(device_index: int):
    if device_index not in _ctxs:
        h = ctypes.c_void_p()
        _lib.call("msr_ctx_create", device_index, ctypes.byref(h))
        _ctxs[device_index] = h
    return _ctxs[device_index]


def _dev_tensor(t, name, dtype=None):
    torch = _torch()
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise TypeError(f"`{name}` must be a torch tensor on a HIP device (the GPU path "
                        "dispatches on device arrays, like the reference's CuArray methods)")
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    return t.contiguous()


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _check_vec(t, n, name, device, dtype):
    """Raises ArgumentError unless t is a contiguous tensor of n elements of
    dtype on device: the library reads and writes n elements through the
    pointer, so a mismatch must stop here (the reference throws
    DimensionMismatch from mul!)."""
    torch = _torch()
    if not isinstance(t, torch.Tensor):
        raise ArgumentError(f"`{name}` must be a tensor")
    if t.device != device:
        raise ArgumentError(f"`{name}` is on {t.device}, the operator on {device}")
    if t.dtype != dtype:
        raise ArgumentError(f"`{name}` has dtype {t.dtype}, expected {dtype}")
    if t.numel() != n:
        raise ArgumentError(f"`{name}` has {t.numel()} elements, expected {n}")
    if not t.is_contiguous():
        raise ArgumentError(f"`{name}` must be contiguous")
    return t


def _check_cmaps(cl, img_shape):
    """Every coil map must be an image of img_shape (row-major view
    img_shape[::-1]); the library reads prod(img_shape) values of each."""
    want = tuple(int(v) for v in img_shape)[::-1]
    for i, c in enumerate(cl):
        if tuple(c.shape) != want:
            raise ArgumentError(f"cmaps[{i}] has shape {tuple(c.shape)}, expected {want} "
                                f"(img_shape {tuple(img_shape)} reversed)")


def _i64arr(vals):
    return (ctypes.c_int64 * len(vals))(*[int(v) for v in vals])


def _cmaps_list(cmaps, img_shape=None):
    torch = _torch()
    if cmaps is None:
        return []
    if isinstance(cmaps, torch.Tensor):
        return [cmaps[i] for i in range(cmaps.shape[0])]
    if isinstance(cmaps, (tuple, list)) and len(cmaps) == 1 and not hasattr(cmaps[0], "shape"):
        return []  # reference default cmaps = (1,)
    return list(cmaps)


def plan_trajectory(trj, sample_mask=None):
    """Device-resident sample selection of a trajectory (masked, frame-major)
    that NFFTNormalOp and calculate_backprojection accept in place of trj, so
    repeated calls skip the selection and the tile sort."""
    return _Traj(trj, sample_mask)


class _Traj:
    """Masked, frame-major sample list on the device (msr_traj)."""

    def __init__(self, trj, sample_mask=None):
        torch = _torch()
        # integer k-space indices select the Cartesian methods (Integer trj
        # dispatch of src/BackProjection.jl:104-156 and FFTNormalOp)
        self.integer = isinstance(trj, torch.Tensor) and not trj.dtype.is_floating_point \
            and not trj.dtype.is_complex
        trj = _dev_tensor(trj, "trj", torch.int32 if self.integer else torch.float32)
        if trj.dim() == 4:  # (Nt, Ncyc, Nr, d): 4D wrapper (src/NFFTNormalOp.jl:45-49)
            trj = trj.reshape(trj.shape[0], -1, trj.shape[3])
        if trj.dim() != 3:
            raise ArgumentError("trj must be (Nt, Nsamp, d) or (Nt, Ncyc, Nr, d)")
        self.nt, self.nsamp, self.d = trj.shape
        self.device = trj.device
        mptr = None
        if sample_mask is not None:
            m = _dev_tensor(sample_mask, "sample_mask")
            m = m.reshape(self.nt, -1)
            if m.shape[1] != self.nsamp:
                raise ArgumentError("sample_mask does not match the trajectory")
            m = m.to(torch.uint8).contiguous()
            self._mask = m
            mptr = _ptr(m)
        self._trj = trj
        h = ctypes.c_void_p()
        _lib.call("msr_traj_create_cartesian" if self.integer else "msr_traj_create",
                  _ctx(self.device.index), self.d, self.nsamp, self.nt,
                  _ptr(trj), mptr, _stream_ptr(self.device), ctypes.byref(h))
        self.h = h
        n = ctypes.c_int64()
        _lib.call("msr_traj_info", h, ctypes.byref(n), None)
        self.M = n.value

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None and _lib._lib is not None:
            _lib.load().msr_traj_destroy(self.h)
            self.h = None


class _NFFTNormalOp:
    """State of the operator (the reference's _NFFTNormalOp struct,
    src/NFFTNormalOp.jl:91-104): ``shape``, ``Ncoeff``, ``Λ`` (packed (Nkm, P)
    device tensor) and ``kmask_indcs`` (1-based int64 device tensor)."""

    def __init__(self, handle, img_shape, device, cmaps_keep):
        self.h = handle
        self.shape = tuple(int(s) for s in img_shape)
        self.device = device
        self._cmaps_keep = cmaps_keep
        K = ctypes.c_int()
        P = ctypes.c_int64()
        nkm = ctypes.c_int64()
        lr = ctypes.c_int()
        _lib.call("msr_normalop_info", handle, ctypes.byref(K), ctypes.byref(P),
                  ctypes.byref(nkm), ctypes.byref(lr))
        self.Ncoeff = K.value
        self.P = P.value
        self.nkm = nkm.value
        self.lambda_real = bool(lr.value)
        cart = ctypes.c_int()
        _lib.call("msr_normalop_kind", handle, ctypes.byref(cart))
        self.cartesian = bool(cart.value)
        self._lam = None
        self._km = None

    def _export(self):
        torch = _torch()
        dt = torch.float32 if self.lambda_real else torch.complex64
        # packed (Nkm, P) for Julia (P, Nkm); Cartesian: (Nkm, K, K) for Julia (K, K, Nkm)
        shape = (self.nkm, self.Ncoeff, self.Ncoeff) if self.cartesian else (self.nkm, self.P)
        lam = torch.empty(shape, dtype=dt, device=self.device)
        km = torch.empty((self.nkm,), dtype=torch.int64, device=self.device)
        _lib.call("msr_normalop_export_kernel", self.h, _ptr(lam), _ptr(km),
                  _stream_ptr(self.device))
        self._lam, self._km = lam, km

    @property
    def Λ(self):
        if self._lam is None:
            self._export()
        return self._lam

    @property
    def kmask_indcs(self):
        if self._km is None:
            self._export()
        return self._km

    def mul_(self, x, b, alpha=1.0, beta=0.0, percoil=False):
        """mul!(x, S, b, α, β) (ext/MRISubspaceReconCUDAExt/NFFTNormalOp.jl:193-218)."""
        torch = _torch()
        n = math.prod(self.shape) * self.Ncoeff
        _check_vec(x, n, "x", self.device, torch.complex64)
        _check_vec(b, n, "b", self.device, torch.complex64)
        a = complex(alpha)
        bt = complex(beta)
        name = "msr_normalop_apply_percoil" if percoil else "msr_normalop_apply"
        _lib.call(name, self.h, _ptr(x), _ptr(b), a.real, a.imag, bt.real, bt.imag,
                  _stream_ptr(self.device))
        return x

    def mul_f64_(self, x, b, alpha=1.0, beta=0.0, compact=False):
        """mul! on ComplexF64 vectors with ComplexF64 arithmetic (the apply of
        the "f64" CG precision; the kernel and coil maps stay ComplexF32).
        compact=True stores the FFT intermediates in 8 bytes per value (the
        apply of the "f64c" precision)."""
        torch = _torch()
        n = math.prod(self.shape) * self.Ncoeff
        _check_vec(x, n, "x", self.device, torch.complex128)
        _check_vec(b, n, "b", self.device, torch.complex128)
        a = complex(alpha)
        bt = complex(beta)
        _lib.call("msr_normalop_apply_f64c" if compact else "msr_normalop_apply_f64", self.h,
                  _ptr(x), _ptr(b), a.real, a.imag, bt.real, bt.imag, _stream_ptr(self.device))
        return x

    def tune_(self, precision="f32", nrep=5):
        """Measured launch configuration of the fused 2D apply for a CG
        precision (the counterpart of plan_fft! with FFTW.MEASURE,
        src/NFFTNormalOp.jl:71-72): every candidate (pass B coil groups
        1/2/4/8, pass A and pass C block orders) runs nrep applies and must
        reproduce the first candidate's result bit for bit; the fastest is
        kept.  Returns {"times": ms per apply per candidate (NaN: not run),
        "chosen": (Gb, order_a, order_c) or None when the operator has no
        fused 2D passes}.  nrep = 0 reports the current choice only."""
        if precision not in _lib.CG_PRECISIONS:
            raise ArgumentError(f"precision must be one of {sorted(_lib.CG_PRECISIONS)}")
        times = (ctypes.c_double * _lib.TUNE_CONFIGS)()
        chosen = ctypes.c_int(-1)
        _lib.call("msr_normalop_tune", self.h, _lib.CG_PRECISIONS[precision], int(nrep), times,
                  ctypes.byref(chosen), _stream_ptr(self.device))
        c = chosen.value
        knobs = None if c < 0 else (1 << (c // 4), (c // 2) % 2, c % 2)
        return {"times": [float(t) for t in times], "chosen": knobs}

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None and _lib._lib is not None:
            _lib.load().msr_normalop_destroy(self.h)
            self.h = None


class _Prod:
    """The ``prod!`` closure of the reference's LinearOperator; ``.A`` is the
    captured _NFFTNormalOp (tests read ``op.prod!.A.Λ``,
    test/reconstruct_radial_gpu.jl:104)."""

    def __init__(self, A):
        self.A = A

    def __call__(self, res, x, alpha=1.0, beta=0.0):
        return self.A.mul_(res, x, alpha, beta)


class LinearOperator:
    """LinearOperator(Complex{T}, n, n, true, true, prod!, nothing, prod!; S)
    (ext/MRISubspaceReconCUDAExt/NFFTNormalOp.jl:59-69)."""

    def __init__(self, A: _NFFTNormalOp):
        self.A = A
        n = math.prod(A.shape) * A.Ncoeff
        self.nrow = self.ncol = n
        self.symmetric = True
        self.hermitian = True
        self.prod = _Prod(A)
        self.tprod = None
        self.ctprod = self.prod

    @property
    def shape(self):
        return (self.nrow, self.ncol)

    def mul_(self, res, v, alpha=1.0, beta=0.0):
        return self.prod(res, v, alpha, beta)

    def __matmul__(self, v):
        torch = _torch()
        v = v.contiguous()
        _check_vec(v, self.ncol, "v", self.A.device, torch.complex64)
        out = torch.empty_like(v)
        return self.prod(out, v, 1.0, 0.0)

    def __repr__(self):
        return (f"Linear operator  nrow: {self.nrow}  ncol: {self.ncol}  eltype: ComplexF32  "
                f"symmetric: true  hermitian: true")


def mul_(res, A, v, alpha=1.0, beta=0.0):
    """LinearAlgebra.mul!(res, A, v, α, β)."""
    if isinstance(A, LinearOperator):
        return A.mul_(res, v, alpha, beta)
    return A.mul_(res, v, alpha, beta)


def _basis(U, traj):
    """U as the library takes it: contiguous (K, Nt) rows (Julia (Nt, K)),
    ComplexF32 or Float32 for a real basis; returns (U, real, K)."""
    torch = _torch()
    U = _dev_tensor(U, "U")
    real = not U.is_complex()
    U = U.to(torch.float32 if real else torch.complex64).contiguous()
    if U.dim() == 1:
        U = U.reshape(1, -1)
    if U.shape[1] != traj.nt:
        raise ArgumentError(f"U has {U.shape[1]} time points, trj {traj.nt}")
    return U, real, U.shape[0]


def _cmaps_arg(cmaps):
    torch = _torch()
    cl = [_dev_tensor(c, "cmaps", torch.complex64) for c in _cmaps_list(cmaps)]
    carr = (ctypes.c_void_p * max(1, len(cl)))(*[c.data_ptr() for c in cl]) if cl else None
    return cl, carr


def NFFTNormalOp(img_shape, a, b, cmaps=None, sample_mask=None, verbose=False):
    """NFFTNormalOp(img_shape, trj, U; cmaps, sample_mask, verbose) or
    NFFTNormalOp(img_shape, Λ, kmask_indcs; cmaps).

    trj/U form: ext/MRISubspaceReconCUDAExt/NFFTNormalOp.jl:1-20 (builds the
    kernel on the device; real float32 U gives the real symmetric kernel).
    Λ/kmask form: :22-70 (Λ packed (Nkm, P), kmask_indcs 1-based int64)."""
    torch = _torch()
    img_shape = tuple(int(s) for s in img_shape)
    d = len(img_shape)
    cl = [_dev_tensor(c, "cmaps", torch.complex64) for c in _cmaps_list(cmaps)]
    _check_cmaps(cl, img_shape)
    carr = (ctypes.c_void_p * max(1, len(cl)))(*[c.data_ptr() for c in cl]) if cl else None
    if isinstance(b, torch.Tensor) and b.dtype in (torch.int64, torch.int32) and b.dim() == 1:
        lam = _dev_tensor(a, "Λ")
        km = _dev_tensor(b, "kmask_indcs", torch.int64)
        if lam.dim() != 2:
            raise ArgumentError("Λ must be packed (Nkm, P)")
        if km.numel() != lam.shape[0]:
            raise AssertionError("length(kmask_indcs) must equal size(Λ, 2) (Julia)")
        real = not lam.is_complex()
        lam = lam.to(torch.float32 if real else torch.complex64).contiguous()
        dev = lam.device
        h = ctypes.c_void_p()
        _lib.call("msr_normalop_create", _ctx(dev.index), d, _i64arr(img_shape), lam.shape[1],
                  int(real), _ptr(lam), _ptr(km), km.numel(), len(cl), carr,
                  _stream_ptr(dev), ctypes.byref(h))
        return LinearOperator(_NFFTNormalOp(h, img_shape, dev, cl))
    traj = a if isinstance(a, _Traj) else _Traj(a, sample_mask)
    U, real, K = _basis(b, traj)
    h = ctypes.c_void_p()
    _lib.call("msr_normalop_create_from_traj", _ctx(traj.device.index), traj.h, d,
              _i64arr(img_shape), K, _ptr(U), int(real), len(cl), carr,
              _stream_ptr(traj.device), ctypes.byref(h))
    return LinearOperator(_NFFTNormalOp(h, img_shape, traj.device, cl))


def FFTNormalOp(a, b=None, c=None, cmaps=None, sample_mask=None, eltype_x=None):
    """FFTNormalOp(img_shape, trj, U; cmaps, sample_mask), FFTNormalOp(M, U; cmaps)
    or FFTNormalOp(Λ; cmaps) (src/FFTNormalOp.jl:28-64), the Cartesian normal
    operator on the device:
      img_shape, trj (Nt, Nsamp, d) integer k-space indices (1..N centred), U (K, Nt);
      M (Nt, *N[::-1]) Bool sampling mask (Julia (N..., Nt)), U (K, Nt);
      Λ (*N[::-1], K, K) (Julia (K, K, N...)) ComplexF32 or Float32.
    A real U gives a real kernel.  ``eltype_x`` is accepted for signature
    parity; vectors are ComplexF32 (ComplexF64 through mul_f64_ / cg)."""
    torch = _torch()
    cl, carr = _cmaps_arg(cmaps)
    h = ctypes.c_void_p()
    if isinstance(a, (tuple, list)):  # (img_shape, trj, U)
        img_shape = tuple(int(v) for v in a)
        traj = b if isinstance(b, _Traj) else _Traj(b, sample_mask)
        if not traj.integer:
            raise ArgumentError("FFTNormalOp takes integer k-space indices (use NFFTNormalOp "
                                "for a non-Cartesian trajectory)")
        U, real, K = _basis(c, traj)
        dev = traj.device
        _check_cmaps(cl, img_shape)
        _lib.call("msr_fftnormalop_create_from_traj", _ctx(dev.index), traj.h, len(img_shape),
                  _i64arr(img_shape), K, _ptr(U), int(real), len(cl), carr, _stream_ptr(dev),
                  ctypes.byref(h))
    elif b is not None:  # (M, U)
        M = _dev_tensor(a, "M")
        nt = M.shape[0]
        img_shape = tuple(int(v) for v in M.shape[1:][::-1])
        M = M.to(torch.uint8).contiguous()
        U = _dev_tensor(b, "U")
        if U.dim() == 1:
            U = U.reshape(1, -1)
        real = not U.is_complex()
        U = U.to(torch.float32 if real else torch.complex64).contiguous()
        if U.shape[1] != nt:
            raise ArgumentError(f"U has {U.shape[1]} time points, M {nt}")
        dev = M.device
        _check_cmaps(cl, img_shape)
        _lib.call("msr_fftnormalop_create_from_mask", _ctx(dev.index), len(img_shape),
                  _i64arr(img_shape), nt, _ptr(M), U.shape[0], _ptr(U), int(real), len(cl), carr,
                  _stream_ptr(dev), ctypes.byref(h))
    else:  # (Λ)
        lam = _dev_tensor(a, "Λ")
        K = lam.shape[-1]
        if lam.shape[-2] != K:
            raise ArgumentError("Λ must be (*N[::-1], K, K)")
        img_shape = tuple(int(v) for v in lam.shape[:-2][::-1])
        real = not lam.is_complex()
        lam = lam.to(torch.float32 if real else torch.complex64).contiguous()
        dev = lam.device
        _check_cmaps(cl, img_shape)
        _lib.call("msr_fftnormalop_create", _ctx(dev.index), len(img_shape), _i64arr(img_shape),
                  K, int(real), _ptr(lam), len(cl), carr, _stream_ptr(dev), ctypes.byref(h))
    return LinearOperator(_NFFTNormalOp(h, img_shape, dev, cl))


def NFFTNormalOp_shared(A, cmaps=None):
    """A further normal operator on A's Toeplitz kernel (shared device memory)
    with its own coil maps, scratch and CG state (msr_normalop_create_shared):
    several reconstructions of one protocol can then be solved at once."""
    torch = _torch()
    src = _op_state(A)
    cl = [_dev_tensor(c, "cmaps", torch.complex64) for c in _cmaps_list(cmaps)]
    _check_cmaps(cl, src.shape)
    carr = (ctypes.c_void_p * max(1, len(cl)))(*[c.data_ptr() for c in cl]) if cl else None
    h = ctypes.c_void_p()
    _lib.call("msr_normalop_create_shared", _ctx(src.device.index), src.h, len(cl), carr,
              _stream_ptr(src.device), ctypes.byref(h))
    return LinearOperator(_NFFTNormalOp(h, src.shape, src.device, cl))


def footprint(A, precision="f64"):
    """(own bytes, kernel bytes) of an operator's device state for a solve at
    `precision` (msr_normalop_footprint)."""
    own = ctypes.c_size_t()
    ker = ctypes.c_size_t()
    _lib.call("msr_normalop_footprint", _op_state(A).h, _prec(precision), ctypes.byref(own),
              ctypes.byref(ker))
    return own.value, ker.value


# ------------------------------------------------ kernel persistence and cache
KERNEL_SOURCES = {0: "built", 1: "memory", 2: "disk"}
_U64x2 = ctypes.c_uint64 * 2


def _key_arg(key):
    if key is None:
        return None
    v = int(key, 16) if isinstance(key, str) else int(key)
    return _U64x2(v >> 64, v & ((1 << 64) - 1))


def _key_hex(k):
    return f"{k[0]:016x}{k[1]:016x}"


def kernel_key(img_shape, trj, U, sample_mask=None):
    """128-bit protocol key (32 hex digits) of the kernel
    NFFTNormalOp(img_shape, trj, U; sample_mask) builds (msr_kernel_key):
    equal for equal image size, basis, masked trajectory and frame sizes."""
    img_shape = tuple(int(s) for s in img_shape)
    traj = trj if isinstance(trj, _Traj) else _Traj(trj, sample_mask)
    U, real, K = _basis(U, traj)
    k = _U64x2()
    _lib.call("msr_kernel_key", _ctx(traj.device.index), traj.h, len(img_shape),
              _i64arr(img_shape), K, _ptr(U), int(real), _stream_ptr(traj.device), k)
    return _key_hex(k)


def save_kernel(A, path, key=None):
    """Writes A's packed kernel and kmask to a kernel file (msr_kernel_save):
    the arrays of the precomputed-kernel constructor
    (ext/MRISubspaceReconCUDAExt/NFFTNormalOp.jl:22-70) with a checksummed header."""
    src = _op_state(A)
    _lib.call("msr_kernel_save", src.h, os.fsencode(path), _key_arg(key), _stream_ptr(src.device))


def load_kernel(path, cmaps=None, key=None, device=None):
    """NFFTNormalOp(img_shape, Λ, kmask_indcs; cmaps) from a kernel file
    (msr_kernel_load); OSError for a missing, truncated or corrupt file,
    ArgumentError if `key` is given and the file holds another protocol."""
    torch = _torch()
    info = kernel_file_info(path, verify=False)
    dev = torch.device("cuda", torch.cuda.current_device() if device is None
                       else torch.device(device).index or 0)
    cl, carr = _cmaps_arg(cmaps)
    h = ctypes.c_void_p()
    _lib.call("msr_kernel_load", _ctx(dev.index), os.fsencode(path), _key_arg(key), len(cl), carr,
              _stream_ptr(dev), ctypes.byref(h))
    return LinearOperator(_NFFTNormalOp(h, info["img_shape"], dev, cl))


def _dev_index(device):
    torch = _torch()
    return torch.cuda.current_device() if device is None else (torch.device(device).index or 0)


def save_tuning(path, device=None):
    """Writes the tuned apply choices measured on the device's context
    (msr_tune_save), the counterpart of FFTW's exported wisdom."""
    _lib.call("msr_tune_save", _ctx(_dev_index(device)), os.fsencode(path))


def load_tuning(path, device=None):
    """Adds a tuning file's choices to the device's context (msr_tune_load):
    operators created afterwards start with them.  Returns the number of
    entries; ArgumentError for a file measured on another device type."""
    n = ctypes.c_int()
    _lib.call("msr_tune_load", _ctx(_dev_index(device)), os.fsencode(path), ctypes.byref(n))
    return n.value


def kernel_file_info(path, verify=True):
    """Header of a kernel file (msr_kernel_file_info; host only): img_shape,
    Ncoeff, lambda_real, nkm, key; verify also checks the payload checksum."""
    d = ctypes.c_int()
    N = (ctypes.c_int64 * 3)()
    K = ctypes.c_int()
    lr = ctypes.c_int()
    nkm = ctypes.c_int64()
    k = _U64x2()
    _lib.call("msr_kernel_file_info", os.fsencode(path), int(bool(verify)), ctypes.byref(d), N,
              ctypes.byref(K), ctypes.byref(lr), ctypes.byref(nkm), k)
    return {"img_shape": tuple(N[: d.value]), "Ncoeff": K.value, "lambda_real": bool(lr.value),
            "nkm": nkm.value, "key": _key_hex(k)}


class KernelCache:
    """Kernels by protocol key (msr_kernel_cache_*): in device memory within
    `budget_bytes` (least recently used dropped first) and, with `directory`,
    in kernel files there.  get() returns a new operator on the protocol's
    kernel, built only on a miss; ``last_source`` says where the kernel came
    from ("built", "memory" or "disk") and ``last_key`` its key."""

    def __init__(self, budget_bytes=32 << 30, directory=None, device=None):
        torch = _torch()
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None
                                   else torch.device(device).index or 0)
        h = ctypes.c_void_p()
        _lib.call("msr_kernel_cache_create", _ctx(self.device.index), int(budget_bytes),
                  None if directory is None else os.fsencode(directory), ctypes.byref(h))
        self.h = h
        self.last_source = None
        self.last_key = None

    def get(self, img_shape, trj, U, cmaps=None, sample_mask=None):
        img_shape = tuple(int(s) for s in img_shape)
        traj = trj if isinstance(trj, _Traj) else _Traj(trj, sample_mask)
        U, real, K = _basis(U, traj)
        cl, carr = _cmaps_arg(cmaps)
        h = ctypes.c_void_p()
        src = ctypes.c_int()
        k = _U64x2()
        _lib.call("msr_kernel_cache_get", self.h, traj.h, len(img_shape), _i64arr(img_shape), K,
                  _ptr(U), int(real), len(cl), carr, _stream_ptr(traj.device), ctypes.byref(h),
                  ctypes.byref(src), k)
        self.last_source = KERNEL_SOURCES[src.value]
        self.last_key = _key_hex(k)
        return LinearOperator(_NFFTNormalOp(h, img_shape, traj.device, cl))

    def stats(self):
        n = ctypes.c_int64()
        b = ctypes.c_size_t()
        c = (ctypes.c_int64 * 3)()
        _lib.call("msr_kernel_cache_stats", self.h, ctypes.byref(n), ctypes.byref(b), c)
        return {"entries": n.value, "bytes": b.value,
                **{KERNEL_SOURCES[i]: c[i] for i in range(3)}}

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None and _lib._lib is not None:
            _lib.load().msr_kernel_cache_destroy(self.h)
            self.h = None


def _test_dimension(data, trj_nt, trj_nsamp, trj_d, U_nt, img_shape, ncoil):
    """test_dimension (src/BackProjection.jl:183-202), on reversed shapes."""
    C, nt, ns = data.shape
    if U_nt != nt:
        raise ArgumentError(f"The second dimension of data ({nt}) and the first one of U "
                            f"({U_nt}) do not match. Both should be number of time points.")
    if trj_d != len(img_shape):
        raise ArgumentError(f"`cmaps` contains {len(img_shape)} image plus one coil dimension, "
                            f"yet the 1ˢᵗ dimension of each trj is of length {trj_d}.")
    if trj_nsamp != ns:
        raise ArgumentError(f"The 2ⁿᵈ dimension of each `trj` is {trj_nsamp} and the 1ˢᵗ "
                            f"dimension of `data` is {ns}.")
    if trj_nt != nt:
        raise ArgumentError(f"`trj` contains {trj_nt} time frames, while data consists of {nt}.")
    if ncoil is not None and ncoil != C:
        raise ArgumentError(f"The last dimension of `cmaps` is {ncoil} and the 3ⁿᵈ dimension "
                            f"of data is {C}.")


def calculate_backprojection(data, trj, cmaps_or_img_shape, U=None, sample_mask=None,
                             density_compensation="none", verbose=False, out=None,
                             channelwise=False):
    """calculate_backprojection(data, trj, cmaps; U, sample_mask,
    density_compensation) -> (K, *N[::-1]) coil-combined
    (ext/MRISubspaceReconCUDAExt/BackProjection.jl:36-71), or with img_shape
    -> (C, K, *N[::-1]) per coil (:1-34).  4D data/trj wrapper :74-79.
    ``trj`` may also be a trajectory handle from ``plan_trajectory`` (reused
    sample selection and tile sort); ``out`` an output tensor to write into.
    ``channelwise=True`` runs the reference's call sequence (one
    single-channel NUFFT per coefficient and coil, :56-66) for timing
    comparisons; it needs U and coil maps."""
    torch = _torch()
    data = _dev_tensor(data, "data", torch.complex64)
    if data.dim() == 4:  # (C, Nt, Ncyc, Nr)
        data = data.reshape(data.shape[0], data.shape[1], -1)
    traj = trj if isinstance(trj, _Traj) else _Traj(trj, sample_mask)
    if traj.integer and not (isinstance(density_compensation, str) and density_compensation == "none"):
        raise ArgumentError("the Cartesian (integer trajectory) backprojection takes no "
                            "density_compensation (src/BackProjection.jl:104,131)")
    per_coil = isinstance(cmaps_or_img_shape, (tuple, list)) and len(cmaps_or_img_shape) > 0 \
        and not hasattr(cmaps_or_img_shape[0], "shape")
    if per_coil:
        img_shape = tuple(int(s) for s in cmaps_or_img_shape)
        cl = []
        ncoil = None
    else:
        cl = [_dev_tensor(c, "cmaps", torch.complex64) for c in _cmaps_list(cmaps_or_img_shape)]
        if not cl:
            raise ArgumentError("calculate_backprojection needs coil maps or img_shape")
        img_shape = tuple(int(s) for s in cl[0].shape[::-1])
        _check_cmaps(cl, img_shape)
        ncoil = len(cl)
    if U is None:
        K = traj.nt
        uptr = None
        real = 0
        u_nt = traj.nt
    else:
        U = _dev_tensor(U, "U")
        if U.dim() == 1:
            U = U.reshape(1, -1)
        real = int(not U.is_complex())
        U = U.to(torch.float32 if real else torch.complex64).contiguous()
        K, u_nt = U.shape
        uptr = _ptr(U)
    _test_dimension(data, traj.nt, traj.nsamp, traj.d, u_nt, img_shape, ncoil)
    C = data.shape[0]
    dcf_ptr = None
    if isinstance(density_compensation, str):
        if density_compensation not in _lib.DCF_MODES:
            raise ErrorException("`density_compensation` can only be `:radial_3D`, "
                                 "`:radial_2D`, `:none`, or of type  `AbstractArray`")
        mode = _lib.DCF_MODES[density_compensation]
    else:
        w = _dev_tensor(torch.as_tensor(density_compensation), "density_compensation",
                        torch.float32).reshape(-1)
        if w.numel() != traj.M:
            raise ArgumentError("density_compensation array must hold one weight per "
                                "masked sample")
        mode = _lib.DCF_ARRAY
        dcf_ptr = _ptr(w)
    shape = ((C, K) if per_coil else (K,)) + img_shape[::-1]
    if out is None:
        out = torch.empty(shape, dtype=torch.complex64, device=data.device)
    elif (out.numel() != math.prod(shape) or out.dtype != torch.complex64
          or not out.is_contiguous() or out.device != data.device):
        raise ArgumentError(f"out must be a contiguous complex64 tensor of {shape}")
    carr = None if per_coil else (ctypes.c_void_p * len(cl))(*[c.data_ptr() for c in cl])
    if channelwise and (per_coil or uptr is None):
        raise ArgumentError("the channel-wise backprojection needs U and coil maps")
    _lib.call("msr_backprojection_channelwise" if channelwise else "msr_backprojection",
              _ctx(data.device.index), traj.h, len(img_shape),
              _i64arr(img_shape), _ptr(data), C, K, uptr, real, carr, mode, dcf_ptr, _ptr(out),
              _stream_ptr(data.device))
    return out


# ------------------------------------------------------ coil-map calibration
def _nextprod235(n):
    m = max(int(n), 1)
    while True:
        r = m
        for p in (2, 3, 5):
            while r % p == 0:
                r //= p
        if r == 1:
            return m
        m += 1


def default_calib_size(img_shape):
    """calib_size default of calculate_coil_maps
    (ext/MRISubspaceReconCUDAExt/CoilMaps.jl:8): nextprod((2,3,5),
    N_i ÷ (maximum(N) ÷ 32))."""
    q = max(img_shape) // 32
    if q == 0:
        raise ZeroDivisionError("maximum(img_shape) ÷ 32 == 0: pass calib_size")
    return tuple(_nextprod235(int(n) // q) for n in img_shape)


def calibration_region(trj, img_shape, calib_size, sample_mask=None):
    """(trj_calib, mask_calib) of calculate_coil_maps
    (ext/MRISubspaceReconCUDAExt/CoilMaps.jl:17-20), on the device
    (msr_calib_select): trj scaled by img_shape ./ calib_size and the samples
    inside the calibration box that the sample mask keeps."""
    torch = _torch()
    trj = _dev_tensor(trj, "trj", torch.float32)
    if trj.dim() == 4:
        trj = trj.reshape(trj.shape[0], -1, trj.shape[3])
    trj = trj.contiguous()
    nt, ns, d = trj.shape
    if len(img_shape) != d or len(calib_size) != d:
        raise ArgumentError("img_shape and calib_size must have one entry per trajectory dimension")
    mptr = None
    if sample_mask is not None:
        m = _dev_tensor(sample_mask, "sample_mask").reshape(nt, ns).to(torch.uint8).contiguous()
        mptr = _ptr(m)
    t_out = torch.empty_like(trj)
    m_out = torch.empty((nt, ns), dtype=torch.uint8, device=trj.device)
    _lib.call("msr_calib_select", _ctx(trj.device.index), d, ns, nt, _ptr(trj), mptr,
              _i64arr(img_shape), _i64arr(calib_size), _ptr(t_out), _ptr(m_out),
              _stream_ptr(trj.device))
    return t_out, m_out.to(torch.bool)


def reconstruct_coilwise(data, trj, img_shape, U=None, sample_mask=None, Niter_cg=100,
                         verbose=False, precision="f64"):
    """reconstruct_coilwise(data, trj, img_shape; U, sample_mask, Niter_cg)
    (ext/MRISubspaceReconCUDAExt/CoilMaps.jl:56-77): one image per coil from
    cg!(x_c, AᴴA, b_c; maxiter=Niter_cg, reltol=0) on the K = 1 operator of
    U[:, 1] (default ones(Nt)); all coils solved in lockstep on one kernel
    (msr_reconstruct_coilwise).  Returns (C, *img_shape[::-1]) complex64.
    precision: CG arithmetic ("f32" is the reference's)."""
    torch = _torch()
    img_shape = tuple(int(s) for s in img_shape)
    data = _dev_tensor(data, "data", torch.complex64)
    if data.dim() == 4:
        data = data.reshape(data.shape[0], data.shape[1], -1)
    data = data.contiguous()
    traj = trj if isinstance(trj, _Traj) else _Traj(trj, sample_mask)
    uptr, real = None, 1
    if U is not None:
        U = _dev_tensor(U, "U")
        if U.dim() == 2:  # (K, Nt): the first Julia column U[:, 1] is row 0
            U = U[0]
        real = int(not U.is_complex())
        U = U.to(torch.float32 if real else torch.complex64).contiguous()
        if U.numel() != traj.nt:
            raise ArgumentError(f"U has {U.numel()} time points, trj {traj.nt}")
        uptr = _ptr(U)
    _test_dimension(data, traj.nt, traj.nsamp, traj.d, traj.nt, img_shape, None)
    C = data.shape[0]
    out = torch.empty((C,) + img_shape[::-1], dtype=torch.complex64, device=data.device)
    _lib.call("msr_reconstruct_coilwise", _ctx(data.device.index), traj.h, len(img_shape),
              _i64arr(img_shape), _ptr(data), C, uptr, real, int(Niter_cg), _prec(precision),
              _ptr(out), _stream_ptr(data.device))
    return out


def calibration_kspace(x):
    """kbp = fftshift(fft(fftshift(x, dims)), dims) over the image axes of
    coil images x (C, *calib_size[::-1]) (ext/MRISubspaceReconCUDAExt/CoilMaps.jl:22-25;
    msr_calib_kspace)."""
    torch = _torch()
    x = _dev_tensor(x, "x", torch.complex64).contiguous()
    d = x.dim() - 1
    if d not in (2, 3):
        raise ArgumentError("x must be (C, Ny, Nx) or (C, Nz, Ny, Nx)")
    out = torch.empty_like(x)
    _lib.call("msr_calib_kspace", _ctx(x.device.index), d, _i64arr(tuple(x.shape[1:][::-1])),
              x.shape[0], _ptr(x), _ptr(out), _stream_ptr(x.device))
    return out


def calculate_coil_maps(data, trj, img_shape, U=None, sample_mask=None, kernel_size=None,
                        calib_size=None, eigThresh_1=0.01, eigThresh_2=0.9, nmaps=1,
                        Niter_cg=100, verbose=False, espirit=None, precision="f64"):
    """calculate_coil_maps (ext/MRISubspaceReconCUDAExt/CoilMaps.jl:1-36): the
    calibration region, the coil-wise reconstruction and the calibration
    k-space run on the device; the ESPIRiT step (MRICoilSensitivities.jl,
    CPU) is the caller's: ``espirit(kbp, img_shape, kernel_size,
    eigThresh_1=..., eigThresh_2=..., nmaps=...)`` receives kbp as a host
    array (calib_size..., C) in Julia order (numpy, column-major) and returns
    maps (img_shape..., C, nmaps).  Without ``espirit`` this raises
    NotImplementedError after nothing has run."""
    img_shape = tuple(int(s) for s in img_shape)
    if espirit is None:
        raise NotImplementedError("ESPIRiT (MRICoilSensitivities.jl) is not part of this library: "
                                  "pass espirit=..., or use calibration_kspace(reconstruct_coilwise(...)) "
                                  "for its input")
    torch = _torch()
    if calib_size is None:
        calib_size = default_calib_size(img_shape)
    calib_size = tuple(int(c) for c in calib_size)
    if kernel_size is None:
        kernel_size = tuple(6 for _ in img_shape)
    traj_src = _dev_tensor(trj, "trj", torch.float32)
    if traj_src.dim() == 4:
        traj_src = traj_src.reshape(traj_src.shape[0], -1, traj_src.shape[3])
    trj_c, mask_c = calibration_region(traj_src, img_shape, calib_size, sample_mask)
    x = reconstruct_coilwise(data, trj_c, calib_size, U=U, sample_mask=mask_c,
                             Niter_cg=Niter_cg, precision=precision)
    kbp = calibration_kspace(x).cpu().numpy().T  # Julia (calib_size..., C) order
    maps = espirit(kbp, img_shape, kernel_size, eigThresh_1=eigThresh_1,
                   eigThresh_2=eigThresh_2, nmaps=nmaps)
    maps = _np().asarray(maps)
    d = len(img_shape)
    if maps.ndim == d + 1:
        maps = maps[..., None]
    # [cmaps[img_idx, ic, in] for ic, in] (:33), coil index fastest
    return [torch.from_numpy(_np().ascontiguousarray(maps[..., c, n].T).astype("complex64"))
            .to(x.device) for n in range(maps.shape[d + 1]) for c in range(maps.shape[d])]


def forward(x, trj, cmaps=None, U=None, sample_mask=None, out=None):
    """Forward subspace model y = A x (the reference tests' data simulation,
    test/reconstruct_radial.jl:62-74): x (K, *N[::-1]) complex64 -> data
    (C, Nt, Nsamp) complex64; samples removed by sample_mask are 0."""
    torch = _torch()
    x = _dev_tensor(x, "x", torch.complex64)
    traj = trj if isinstance(trj, _Traj) else _Traj(trj, sample_mask)
    cl = [_dev_tensor(c, "cmaps", torch.complex64) for c in _cmaps_list(cmaps)]
    img_shape = tuple(int(s) for s in x.shape[1:][::-1])
    if len(img_shape) != traj.d:
        raise ArgumentError(f"x has {len(img_shape)} image dimensions, trj {traj.d}")
    _check_cmaps(cl, img_shape)
    if U is None:
        K, uptr, real = traj.nt, None, 0
    else:
        U = _dev_tensor(U, "U")
        if U.dim() == 1:
            U = U.reshape(1, -1)
        real = int(not U.is_complex())
        U = U.to(torch.float32 if real else torch.complex64).contiguous()
        K = U.shape[0]
        uptr = _ptr(U)
    if x.shape[0] != K:
        raise ArgumentError(f"x has {x.shape[0]} coefficients, U {K}")
    C = max(1, len(cl))
    shape = (C, traj.nt, traj.nsamp)
    if out is None:
        out = torch.empty(shape, dtype=torch.complex64, device=x.device)
    else:
        _check_vec(out, math.prod(shape), "out", x.device, torch.complex64)
    carr = (ctypes.c_void_p * len(cl))(*[c.data_ptr() for c in cl]) if cl else None
    _lib.call("msr_forward", _ctx(x.device.index), traj.h, len(img_shape), _i64arr(img_shape),
              _ptr(x), K, uptr, real, len(cl), carr, _ptr(out), _stream_ptr(x.device))
    return out


REL_TOL_F32 = math.sqrt(2.0 ** -23)
_side = {}


def _side_stream(device):
    torch = _torch()
    if device.index not in _side:
        _side[device.index] = torch.cuda.Stream(device)  # default priority
    return _side[device.index]


def _prec(precision):
    try:
        return _lib.CG_PRECISIONS[precision]
    except KeyError:
        raise ArgumentError(f"precision must be one of {sorted(_lib.CG_PRECISIONS)}, "
                            f"got {precision!r}") from None


def cg_(x, A, b, maxiter=None, reltol=REL_TOL_F32, abstol=0.0, initially_zero=False,
        mode="device", precision="f64", log=False):
    """cg!(x, A, b; maxiter, reltol, abstol) of IterativeSolvers.jl 0.9 on the
    device (FP64 dot products, norms and step lengths).  mode: "device"
    (device-resident scalars), "graph" (hipGraph-captured iterations) or
    "host" (scalars read back every iteration, per-coil applies: the plain
    reference-order GPU path).  precision: "f64" (FP64 vectors and a
    ComplexF64 apply), "f64c" (the same with the 2D apply's FFT intermediates
    stored in 8 bytes per value), "mixed" (FP64 vectors, ComplexF32 apply) or
    "f32" (ComplexF32 vectors and apply, the reference's arithmetic).  x and b
    are ComplexF32 in every precision."""
    op = A.A if isinstance(A, LinearOperator) else A
    torch = _torch()
    n = math.prod(op.shape) * op.Ncoeff
    _check_vec(x, n, "x", op.device, torch.complex64)
    _check_vec(b, n, "b", op.device, torch.complex64)
    if maxiter is None:
        maxiter = n
    it = ctypes.c_int()
    res = ctypes.c_double()
    cur = torch.cuda.current_stream(op.device)
    side = None
    if mode == "graph" and cur.cuda_stream == 0:
        # stream capture needs a non-default stream: run on a side stream
        # ordered after (and before) the caller's stream
        side = _side_stream(op.device)
        side.wait_stream(cur)
    st = ctypes.c_void_p((side or cur).cuda_stream)
    # the iteration count and residual are read back (a stream synchronisation)
    # only when asked for
    _lib.call("msr_cg", op.h, _ptr(x), _ptr(b), int(maxiter), float(reltol), float(abstol),
              int(bool(initially_zero)), _lib.CG_MODES[mode], _prec(precision),
              ctypes.byref(it) if log else None, ctypes.byref(res) if log else None, st)
    if side is not None:
        cur.wait_stream(side)
    if log:
        return x, {"iters": it.value, "residual": res.value}
    return x


def cg(A, b, **kw):
    """cg(A, b; maxiter, ...) = cg!(zeros, A, b; initially_zero=true)."""
    torch = _torch()
    b = b.contiguous()
    x = torch.empty_like(b)
    return cg_(x, A, b, initially_zero=True, **kw)


def _op_state(A):
    return A.A if isinstance(A, LinearOperator) else A


class NFFTNormalOpBatch:
    """Many normal operators (any mix of image size, coil count, K, 2D/3D)
    applied and solved together (msr_batch_*): every pass of the fused apply is
    one launch per group of operators sharing its FFT length, and the CG
    vector steps are one launch for the whole batch.  Each operator's result
    is the one its own mul!/cg! gives.  The batch borrows the operators (they
    must stay alive and must not be used elsewhere during a batch call)."""

    def __init__(self, ops):
        ops = [_op_state(A) for A in ops]
        if not ops:
            raise ArgumentError("a batch needs at least one operator")
        dev = ops[0].device
        if any(A.device != dev for A in ops):
            raise ArgumentError("the operators of a batch must live on one device")
        self.ops = ops
        self.device = dev
        arr = (ctypes.c_void_p * len(ops))(*[A.h.value for A in ops])
        h = ctypes.c_void_p()
        _lib.call("msr_batch_create", _ctx(dev.index), len(ops), arr, ctypes.byref(h))
        self.h = h

    def __len__(self):
        return len(self.ops)

    def _ptrs(self, ts, name):
        if len(ts) != len(self.ops):
            raise ArgumentError(f"{name} has {len(ts)} vectors for {len(self.ops)} operators")
        torch = _torch()
        for i, (t, A) in enumerate(zip(ts, self.ops)):
            _check_vec(t, math.prod(A.shape) * A.Ncoeff, f"{name}[{i}]", self.device,
                       torch.complex64)
        return (ctypes.c_void_p * len(ts))(*[t.data_ptr() for t in ts])

    def mul_(self, xs, bs, alpha=1.0, beta=0.0):
        """x_j = β x_j + α A_j b_j for every operator."""
        a, bt = complex(alpha), complex(beta)
        _lib.call("msr_batch_apply", self.h, self._ptrs(xs, "x"), self._ptrs(bs, "b"),
                  a.real, a.imag, bt.real, bt.imag, _stream_ptr(self.device))
        return xs

    def cg_(self, xs, bs, maxiter, reltol=None, abstol=0.0, initially_zero=False,
            mode="device", precision="f64", log=False):
        """cg!(x_j, A_j, b_j) for every operator in lockstep (modes "device"
        and "graph"); see cg_ for the recurrence and the precision."""
        reltol = REL_TOL_F32 if reltol is None else reltol
        torch = _torch()
        cur = torch.cuda.current_stream(self.device)
        side = None
        if mode == "graph" and cur.cuda_stream == 0:
            side = _side_stream(self.device)
            side.wait_stream(cur)
        st = ctypes.c_void_p((side or cur).cuda_stream)
        n = len(self.ops)
        its = (ctypes.c_int * n)()
        res = (ctypes.c_double * n)()
        _lib.call("msr_batch_cg", self.h, self._ptrs(xs, "x"), self._ptrs(bs, "b"), int(maxiter),
                  float(reltol), float(abstol), int(bool(initially_zero)), _lib.CG_MODES[mode],
                  _prec(precision), its if log else None, res if log else None, st)
        if side is not None:
            cur.wait_stream(side)
        if log:
            return xs, {"iters": list(its), "residual": list(res)}
        return xs

    # stepwise solve (continuous batching)
    def cg_start(self, xs, bs, maxiter, reltol=None, abstol=0.0, initially_zero=True,
                 precision="f64", stream=None):
        reltol = REL_TOL_F32 if reltol is None else reltol
        _lib.call("msr_batch_cg_start", self.h, self._ptrs(xs, "x"), self._ptrs(bs, "b"),
                  int(maxiter), float(reltol), float(abstol), int(bool(initially_zero)),
                  _prec(precision), self._st(stream))

    def cg_iterate(self, niter, stream=None):
        _lib.call("msr_batch_cg_iterate", self.h, int(niter), self._st(stream))

    def cg_finish(self, stream=None):
        _lib.call("msr_batch_cg_finish", self.h, self._st(stream))

    def cg_status(self, stream=None):
        n = len(self.ops)
        its = (ctypes.c_int * n)()
        res = (ctypes.c_double * n)()
        done = (ctypes.c_int * n)()
        _lib.call("msr_batch_cg_status", self.h, its, res, done, self._st(stream))
        return list(its), list(res), [bool(d) for d in done]

    def _st(self, stream):
        if stream is None:
            return _stream_ptr(self.device)
        return ctypes.c_void_p(stream.cuda_stream)

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None and _lib._lib is not None:
            _lib.load().msr_batch_destroy(self.h)
            self.h = None


__all__ += ["NFFTNormalOpBatch"]


class NFFTServer:
    """Continuous batching of CG solves (msr_server_*): jobs wait in a queue
    (admitted in arrival order, or least estimated work first with
    ``policy="shortest_first"``), are admitted while the device state of the jobs in flight fits
    ``budget_bytes`` (at most ``max_active`` of them), run in lockstep batched
    CG iterations and retire when their CG stops.  Each job's result is the
    one ``cg_`` gives for its operator alone (same precision and stopping)."""

    QUEUED, RUNNING, DONE = 0, 1, 2
    POLICIES = {"fifo": 0, "shortest_first": 1}

    def __init__(self, device=None, budget_bytes=None, max_active=64, precision="f64",
                 policy="fifo"):
        torch = _torch()
        dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        if budget_bytes is None:
            free, _total = torch.cuda.mem_get_info(dev)
            budget_bytes = int(free * 0.8)
        self.device = dev
        self.precision = precision
        self._keep = {}
        h = ctypes.c_void_p()
        _lib.call("msr_server_create", _ctx(dev.index), ctypes.c_size_t(int(budget_bytes)),
                  int(max_active), _prec(precision), ctypes.byref(h))
        self.h = h
        if policy not in self.POLICIES:
            raise ArgumentError(f"server policy {policy!r}: one of {sorted(self.POLICIES)}")
        _lib.call("msr_server_set_policy", h, self.POLICIES[policy])

    def submit(self, A, x, b, maxiter, reltol=0.0, abstol=0.0):
        """Queue the solve of A x = b (x, b: contiguous complex64 device vectors
        of length V*K); returns the job id."""
        op = _op_state(A)
        n = math.prod(op.shape) * op.Ncoeff
        torch = _torch()
        for t, name in ((x, "x"), (b, "b")):
            _check_vec(t, n, name, self.device, torch.complex64)
        jid = ctypes.c_int64()
        _lib.call("msr_server_submit", self.h, op.h, _ptr(x), _ptr(b), int(maxiter), float(reltol),
                  float(abstol), ctypes.byref(jid))
        self._keep[jid.value] = (A, x, b)
        return jid.value

    def step(self, niter=5, stream=None):
        """Admit, run niter batched CG iterations, retire; returns
        (jobs in flight, jobs queued)."""
        na = ctypes.c_int()
        nq = ctypes.c_int()
        st = _stream_ptr(self.device) if stream is None else ctypes.c_void_p(stream.cuda_stream)
        _lib.call("msr_server_step", self.h, int(niter), ctypes.byref(na), ctypes.byref(nq), st)
        return na.value, nq.value

    def poll(self, job):
        """{'state', 'iters', 'residual', 'queued_ms', 'running_ms', 'total_ms'}."""
        state = ctypes.c_int()
        its = ctypes.c_int()
        res = ctypes.c_double()
        t = (ctypes.c_double * 3)()
        _lib.call("msr_server_poll", self.h, int(job), ctypes.byref(state), ctypes.byref(its),
                  ctypes.byref(res), t)
        return {"state": state.value, "iters": its.value, "residual": res.value,
                "queued_ms": t[0], "running_ms": t[1], "total_ms": t[2]}

    def forget(self, job):
        _lib.call("msr_server_forget", self.h, int(job))
        self._keep.pop(job, None)

    def run(self, niter=5):
        """Steps until no job is in flight or queued."""
        while True:
            na, nq = self.step(niter)
            if na == 0 and nq == 0:
                return

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None and _lib._lib is not None:
            _lib.load().msr_server_destroy(self.h)
            self.h = None
