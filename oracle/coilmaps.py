"""FP64 restatement of the GPU coil-map calibration front end (test
infrastructure only): the calibration-region selection, the coil-wise
reconstruction at K = 1 and the calibration k-space handed to ESPIRiT.
ESPIRiT itself (MRICoilSensitivities.jl 0.1.1, a third-party CPU package,
Project.toml:8,19) is outside the path and not restated.

Layout as in oracle.toeplitz: trj (Nt, Nsamp, d), sample_mask (Nt, Nsamp),
data (C, Nt, Nsamp), U (K, Nt), images (C, *N[::-1]).
"""
from __future__ import annotations

import numpy as np

from . import toeplitz
from .cg import cg


def nextprod235(n: int) -> int:
    """Smallest 2^a 3^b 5^c >= n (Julia nextprod((2, 3, 5), n))."""
    m = max(int(n), 1)
    while True:
        r = m
        for p in (2, 3, 5):
            while r % p == 0:
                r //= p
        if r == 1:
            return m
        m += 1


def calib_size_default(img_shape):
    """calib_size default of calculate_coil_maps
    (ext/MRISubspaceReconCUDAExt/CoilMaps.jl:8):
    nextprod((2,3,5), N_i ÷ (maximum(N) ÷ 32))."""
    q = max(img_shape) // 32
    if q == 0:
        raise ZeroDivisionError("maximum(img_shape) ÷ 32 == 0")
    return tuple(nextprod235(n // q) for n in img_shape)


def calib_select(trj, img_shape, calib_size, sample_mask=None):
    """(trj_calib, mask_calib) of calculate_coil_maps
    (ext/MRISubspaceReconCUDAExt/CoilMaps.jl:17-20): the scale N ./ calib_size
    rounded to Float32 (cu of a Float64 array), a sample kept if every
    |k_a| * scale_a < 0.5 (in Float32) and the sample mask keeps it, and the
    trajectory multiplied by the scale (in Float32)."""
    trj = np.asarray(trj, dtype=np.float32)
    d = trj.shape[-1]
    scale = np.array([img_shape[a] / calib_size[a] for a in range(d)], dtype=np.float64)
    scale = scale.astype(np.float32)
    inside = np.all(np.abs(trj) * scale < 0.5, axis=-1)
    if sample_mask is not None:
        inside &= np.asarray(sample_mask, dtype=bool).reshape(inside.shape)
    return (trj * scale).astype(np.float32), inside


def reconstruct_coilwise(data, trj, img_shape, U=None, sample_mask=None, niter=100):
    """reconstruct_coilwise (ext/MRISubspaceReconCUDAExt/CoilMaps.jl:56-77;
    CPU form src/CoilMaps.jl:118-139): the K = 1 operator of U[:, 1]
    (default ones(Nt)), the per-coil backprojection with U[:, 1], and
    cg!(x_c, AᴴA, b_c; maxiter=niter, reltol=0) per coil.  Returns
    (C, *N[::-1]) complex128."""
    data = np.asarray(data)
    C, nt, _ = data.shape
    u1 = np.ones((1, nt), dtype=np.float32) if U is None else np.asarray(U)[:1]
    Lp, km = toeplitz.toeplitz_kernel(img_shape, trj, u1, sample_mask)
    op = toeplitz.NormalOp(img_shape, Lp, km, None)
    xbp = toeplitz.backprojection(data, trj, tuple(img_shape), U=u1, sample_mask=sample_mask)
    xbp = xbp.reshape((C,) + tuple(img_shape)[::-1])
    out = np.empty(xbp.shape, dtype=np.complex128)
    for c in range(C):
        out[c] = cg(op, xbp[c].reshape(-1), niter, reltol=0.0).reshape(out.shape[1:])
    return out


def calib_kspace(x, d):
    """kbp = fftshift(fft(fftshift(x, dims)), dims) over the d image axes of
    each coil image (ext/MRISubspaceReconCUDAExt/CoilMaps.jl:22-25)."""
    axes = tuple(range(-d, 0))
    x = np.asarray(x, dtype=np.complex128)
    return np.fft.fftshift(np.fft.fftn(np.fft.fftshift(x, axes=axes), axes=axes), axes=axes)
