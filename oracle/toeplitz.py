"""FP64 restatement of the Toeplitz subspace normal operator, its kernel and
the subspace backprojection (test infrastructure only).

Layout: arrays are the row-major view of the reference's Julia arrays (shapes
reversed): trj (Nt, Nsamp, d), sample_mask (Nt, Nsamp), data (C, Nt, Nsamp),
U (K, Nt), images (K, *N[::-1]), cmaps (C, *N[::-1]), packed kernel
(Nkm, P) for Julia (P, Nkm), full kernel (Nkm, K, K) for Julia (K, K, Nkm)
indexed [i, col, row] (so Lam_full[i, k, j] = Julia Lam[j, k, i]).
"""
from __future__ import annotations

import math

import numpy as np
import scipy.fft as sfft

from . import nufft


# --------------------------------------------------------------------------
# sample selection (src/NFFTNormalOp.jl:126-131, src/BackProjection.jl:26-31)
# --------------------------------------------------------------------------
def select_samples(trj, sample_mask=None):
    """trj[:, mask] of the reference: frame-major list of masked samples.
    Returns (kc (M, d) float32, frame (M,) int64, nsamp_t (Nt,))."""
    trj = np.asarray(trj)
    nt, ns, d = trj.shape
    if sample_mask is None:
        sample_mask = np.ones((nt, ns), dtype=bool)
    sample_mask = np.asarray(sample_mask, dtype=bool).reshape(nt, ns)
    nsamp_t = sample_mask.sum(axis=1)
    if nsamp_t.sum() == 0:
        raise AssertionError("Mask removes all samples")
    kc = trj[sample_mask]
    frame = np.repeat(np.arange(nt), nsamp_t)
    return kc, frame, nsamp_t


def packed_index(j, k):
    """0-based packed index of (j, k), j <= k: LAPACK 'U' packing
    (ext/MRISubspaceReconCUDAExt/NFFTNormalOp.jl:239, 1-based j + k(k-1)/2)."""
    return j + k * (k + 1) // 2


def unpack(Lp, K):
    """Packed (Nkm, P) -> full (Nkm, K, K) with Lam_full[:, k, j] the Julia
    Lam[j, k] (column k, row j): upper triangle from packed storage, lower
    triangle conj (ext/.../NFFTNormalOp.jl:255-262) or the same real value
    (:278-281)."""
    nkm = Lp.shape[0]
    real = not np.iscomplexobj(Lp)
    full = np.empty((nkm, K, K), dtype=np.complex128)
    for k in range(K):
        for j in range(K):
            if j <= k:
                full[:, k, j] = Lp[:, packed_index(j, k)]
            else:
                v = Lp[:, packed_index(k, j)]
                full[:, k, j] = v if real else np.conj(v)
    return full


# --------------------------------------------------------------------------
# kmask (src/NFFTNormalOp.jl:106-118; ext/.../NFFTNormalOp.jl:75-86)
# --------------------------------------------------------------------------
KMASK_HALF_SUPPORT = 4


def kmask_cells(kc, n):
    """Base cell of each sample on a grid of size n per dim, as NonuniformFFTs'
    spreading locates x = mod(-2 pi k, 2 pi) on cells of width 2 pi / n:
    floor(mod(-k, 1) * n), evaluated exactly in float64 from float32 k."""
    kc = np.asarray(kc, dtype=np.float32).astype(np.float64)
    t = -kc
    t = t - np.floor(t)
    nn = np.asarray(n, dtype=np.float64)[None, :]
    cell = np.floor(t * nn).astype(np.int64)
    cell = np.minimum(cell, np.asarray(n, dtype=np.int64)[None, :] - 1)
    return cell


def kmask_indcs(n, kc):
    """Sorted 0-based linear indices (x fastest) of the cells of the grid n
    (= 2 .* img_shape) reached by a half-support-4 spreading of every sample:
    cells base-3 .. base+4 per dim, periodic."""
    n = tuple(int(v) for v in n)
    d = len(n)
    for v in n:
        m = v
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m != 1:
            raise AssertionError("img_shape_os has to be composed of the prime factors 2, 3, and 5")
    cell = kmask_cells(kc, n)
    hs = KMASK_HALF_SUPPORT
    offs = np.arange(-hs + 1, hs + 1)
    marks = np.zeros(n[::-1], dtype=bool)
    # mark per dim combinations (vectorised over samples)
    if d == 2:
        ix = np.mod(cell[:, 0, None] + offs[None], n[0])  # (M, 8)
        iy = np.mod(cell[:, 1, None] + offs[None], n[1])
        marks[iy[:, :, None], ix[:, None, :]] = True
    else:
        # chunks of samples keep the (m, 8, 8, 8) index arrays small
        for a in range(0, cell.shape[0], 65536):
            cs = cell[a:a + 65536]
            ix = np.mod(cs[:, 0, None] + offs[None], n[0])
            iy = np.mod(cs[:, 1, None] + offs[None], n[1])
            iz = np.mod(cs[:, 2, None] + offs[None], n[2])
            marks[iz[:, :, None, None], iy[:, None, :, None], ix[:, None, None, :]] = True
    return np.flatnonzero(marks.reshape(-1)).astype(np.int64)


# --------------------------------------------------------------------------
# Toeplitz kernel (src/NFFTNormalOp.jl:121-225; ext/.../NFFTNormalOp.jl:89-191)
# --------------------------------------------------------------------------
def psf_grid(kc, frame, U, n, j, k):
    """p[r] = sum_m conj(U[t_m, j]) U[t_m, k] exp(+2 pi i k_m . r) on the grid n
    in FFT order (src/NFFTNormalOp.jl:151-157), FP64."""
    U = np.asarray(U)
    if np.iscomplexobj(U):
        S = np.conj(U[j, frame].astype(np.complex128)) * U[k, frame].astype(np.complex128)
    else:
        S = U[j, frame].astype(np.float64) * U[k, frame].astype(np.float64)
    return nufft.type1(kc, S[None].astype(np.complex128), n, shift=False)[0]


def toeplitz_kernel(img_shape, trj, U, sample_mask=None, kmask=None):
    """Packed kernel (Nkm, P) and kmask (0-based) for the 2N grid.

    Complex U: Lp[:, packed(j,k)] = fftn(p_jk)[kmask] for j <= k
    (src/NFFTNormalOp.jl:148-163, ext/.../NFFTNormalOp.jl:123-135).
    Real U: Lp real = Re(fftn(p_jk))[kmask] (see module header of oracle)."""
    n = tuple(2 * int(v) for v in img_shape)
    kc, frame, _ = select_samples(trj, sample_mask)
    if kmask is None:
        kmask = kmask_indcs(n, kc)
    U = np.asarray(U)
    K = U.shape[0]
    P = K * (K + 1) // 2
    real = not np.iscomplexobj(U)
    Lp = np.empty((kmask.size, P), dtype=np.float64 if real else np.complex128)
    for k in range(K):
        for j in range(k + 1):
            p = psf_grid(kc, frame, U, n, j, k)
            lam = sfft.fftn(p, workers=-1).reshape(-1)
            Lp[:, packed_index(j, k)] = lam[kmask].real if real else lam[kmask]
    return Lp, kmask


def dense_kernel(img_shape, Lp, kmask, K):
    """Full kernel on the whole grid with zeros off the mask:
    (K, K, *n[::-1]) with [k, j] = Julia Lam[j, k]."""
    n = tuple(2 * int(v) for v in img_shape)
    G = int(np.prod(n))
    full = unpack(np.asarray(Lp), K)
    dense = np.zeros((K, K, G), dtype=np.complex128)
    dense[:, :, kmask] = np.transpose(full, (1, 2, 0))
    return dense.reshape((K, K) + n[::-1])


# --------------------------------------------------------------------------
# Normal-operator apply (src/NFFTNormalOp.jl:227-267; ext/.../NFFTNormalOp.jl:193-218)
# --------------------------------------------------------------------------
def apply_normal(img_shape, Lp, kmask, cmaps, b, alpha=1.0, beta=0.0, x=None):
    """x <- beta*x + alpha * sum_c S_c^H Crop IFFT(Lam . FFT(Pad(S_c b))).

    beta == 0 overwrites x (src/NFFTNormalOp.jl:232-236).  cmaps None means the
    reference's default cmaps=(1,).  b, x: (K, *N[::-1])."""
    img_shape = tuple(int(v) for v in img_shape)
    d = len(img_shape)
    n = tuple(2 * v for v in img_shape)
    K = b.shape[0]
    dense = dense_kernel(img_shape, Lp, kmask, K)
    b = np.asarray(b, dtype=np.complex128)
    if x is None or beta == 0:
        out = np.zeros_like(b)
    else:
        out = beta * np.asarray(x, dtype=np.complex128)
    maps = [None] if cmaps is None else list(np.asarray(cmaps, dtype=np.complex128))
    crop = tuple(slice(0, v) for v in img_shape[::-1])
    axes = tuple(range(1, d + 1))
    for cm in maps:
        kl1 = np.zeros((K,) + n[::-1], dtype=np.complex128)
        kl1[(slice(None),) + crop] = b if cm is None else cm[None] * b
        kl1 = sfft.fftn(kl1, axes=axes, workers=-1)
        kl2 = np.einsum("kj...,k...->j...", dense, kl1)
        img = sfft.ifftn(kl2, axes=axes, workers=-1)[(slice(None),) + crop]
        out += alpha * (img if cm is None else np.conj(cm)[None] * img)
    return out


class NormalOp:
    """apply_normal with the dense kernel built once (for repeated applies)."""

    def __init__(self, img_shape, Lp, kmask, cmaps):
        self.N = tuple(int(v) for v in img_shape)
        self.n = tuple(2 * v for v in self.N)
        self.K = int((np.sqrt(8 * np.asarray(Lp).shape[1] + 1) - 1) // 2)
        self.dense = dense_kernel(self.N, Lp, kmask, self.K)
        self.maps = [None] if cmaps is None else list(np.asarray(cmaps, dtype=np.complex128))

    def __call__(self, v):
        K = self.K
        d = len(self.N)
        b = np.asarray(v, dtype=np.complex128).reshape((K,) + self.N[::-1])
        out = np.zeros_like(b)
        crop = tuple(slice(0, w) for w in self.N[::-1])
        axes = tuple(range(1, d + 1))
        for cm in self.maps:
            kl1 = np.zeros((K,) + self.n[::-1], dtype=np.complex128)
            kl1[(slice(None),) + crop] = b if cm is None else cm[None] * b
            kl1 = sfft.fftn(kl1, axes=axes, workers=-1)
            kl2 = np.einsum("kj...,k...->j...", self.dense, kl1)
            img = sfft.ifftn(kl2, axes=axes, workers=-1)[(slice(None),) + crop]
            out += img if cm is None else np.conj(cm)[None] * img
        return out.reshape(-1)


# --------------------------------------------------------------------------
# Backprojection (src/BackProjection.jl:23-102,169-181; ext/.../BackProjection.jl)
# --------------------------------------------------------------------------
def density_weights(kc, density_compensation):
    """apply_density_compensation! (src/BackProjection.jl:169-181)."""
    kc = np.asarray(kc, dtype=np.float64)
    if isinstance(density_compensation, str):
        if density_compensation == "radial_3D":
            return np.sum(kc ** 2, axis=1)
        if density_compensation == "radial_2D":
            return np.sqrt(np.sum(kc ** 2, axis=1))
        if density_compensation == "none":
            return None
        raise ValueError("`density_compensation` can only be `:radial_3D`, `:radial_2D`, "
                         "`:none`, or of type  `AbstractArray`")
    return np.asarray(density_compensation, dtype=np.float64).reshape(-1)


def backprojection(data, trj, cmaps_or_shape, U=None, sample_mask=None,
                   density_compensation="none"):
    """Coil-combined (cmaps given): (K, *N[::-1]),
    b_k = sum_c conj(S_c) * type1(w * y_c * conj(U[t, k]))  (src/BackProjection.jl:62-102).
    Per-coil (img_shape given): (C, K, *N[::-1])  (src/BackProjection.jl:23-60)."""
    data = np.asarray(data)
    C, nt, ns = data.shape
    kc, frame, _ = select_samples(trj, sample_mask)
    if U is None:
        U = np.eye(nt, dtype=np.complex128)  # I(Nt): one coefficient per frame
    U = np.asarray(U)
    K = U.shape[0]
    if sample_mask is None:
        yr = data.reshape(C, -1).astype(np.complex128)
    else:
        yr = data[:, np.asarray(sample_mask, dtype=bool).reshape(nt, ns)].astype(np.complex128)
    w = density_weights(kc, density_compensation)
    if w is not None:
        yr = yr * w[None]
    Uc = np.conj(U[:, frame].astype(np.complex128))  # (K, M)
    v = (yr[:, None, :] * Uc[None, :, :]).reshape(C * K, -1)
    if isinstance(cmaps_or_shape, (tuple, list)) and np.isscalar(cmaps_or_shape[0]):
        img_shape = tuple(int(s) for s in cmaps_or_shape)
        out = nufft.type1(kc, v, img_shape, shift=True)
        return out.reshape((C, K) + img_shape[::-1])
    cmaps = np.asarray(cmaps_or_shape, dtype=np.complex128)
    img_shape = cmaps.shape[1:][::-1]
    out = nufft.type1(kc, v, img_shape, shift=True).reshape((C, K) + img_shape[::-1])
    return np.sum(np.conj(cmaps)[:, None] * out, axis=0)


def forward(x, trj, U, cmaps, sample_mask=None):
    """Forward model y_{m,c} = sum_k U[t_m, k] type2(S_c x_k)(k_m) used by the
    reference tests to simulate data (test/reconstruct_radial.jl:62-74).
    Returns data (C, Nt, Nsamp); masked-out samples are zero."""
    trj = np.asarray(trj)
    nt, ns, d = trj.shape
    kc, frame, _ = select_samples(trj, sample_mask)
    U = np.asarray(U)
    K = U.shape[0]
    cm = np.asarray(cmaps, dtype=np.complex128)
    C = cm.shape[0]
    xs = (cm[:, None] * np.asarray(x, dtype=np.complex128)[None]).reshape((C * K,) + x.shape[1:])
    yk = nufft.type2(kc, xs, shift=True).reshape(C, K, -1)
    y = np.einsum("ckm,km->cm", yk, U[:, frame].astype(np.complex128))
    out = np.zeros((C, nt, ns), dtype=np.complex128)
    if sample_mask is None:
        out[:] = y.reshape(C, nt, ns)
    else:
        out[:, np.asarray(sample_mask, dtype=bool).reshape(nt, ns)] = y
    return out
