"""FP64 restatement of the Cartesian normal operator and backprojection (test
infrastructure only): src/FFTNormalOp.jl and the Integer-trajectory methods
of src/BackProjection.jl:104-156.

Layout as in oracle.toeplitz (row-major views of the Julia arrays): integer
trj (Nt, Nsamp, d), sample_mask (Nt, Nsamp), mask M (Nt, *N[::-1]), data
(C, Nt, Nsamp), U (K, Nt), images (K, *N[::-1]), cmaps (C, *N[::-1]); the
kernel is kept as Lam (G, K, K) over the cells q of the N grid (x fastest),
Lam[q, j, k] = Julia Λ[j, k, q].
"""
from __future__ import annotations

import numpy as np


def sample_cells(trj, img_shape):
    """Cell of each sample: k_idx = mod1(trj - N ÷ 2, N) per axis (the
    ifftshift of src/FFTNormalOp.jl:89 and src/BackProjection.jl:116), as a
    0-based linear index with x fastest.  trj (..., d) integers."""
    trj = np.asarray(trj, dtype=np.int64)
    q = np.zeros(trj.shape[:-1], dtype=np.int64)
    stride = 1
    for a, n in enumerate(img_shape):
        q += np.mod(trj[..., a] - n // 2 - 1, n) * stride
        stride *= n
    return q


def kernel_cartesian(img_shape, trj, U, sample_mask=None):
    """calculate_kernel_cartesian(img_shape, trj, U; sample_mask)
    (src/FFTNormalOp.jl:81-99): Λ[j, k, k_idx] += conj(U[t, j]) U[t, k] for
    every kept sample of frame t.  Returns Lam (G, K, K) complex128 (float64
    for a real U)."""
    trj = np.asarray(trj)
    nt, ns, d = trj.shape
    U = np.asarray(U)
    K = U.shape[0]
    G = int(np.prod(img_shape))
    real = not np.iscomplexobj(U)
    U64 = U.astype(np.float64 if real else np.complex128)
    q = sample_cells(trj, img_shape)                      # (Nt, Nsamp)
    keep = np.ones((nt, ns), bool) if sample_mask is None else np.asarray(sample_mask, bool).reshape(nt, ns)
    counts = np.zeros((G, nt))
    for t in range(nt):
        np.add.at(counts[:, t], q[t][keep[t]], 1.0)
    # Λ[q, j, k] = sum_t counts[q, t] conj(U[j, t]) U[k, t]
    return np.einsum("qt,jt,kt->qjk", counts, np.conj(U64), U64)


def kernel_cartesian_mask(M, U):
    """calculate_kernel_cartesian(M, U) (src/FFTNormalOp.jl:101-111): M
    (Nt, *N[::-1]) binary is ifftshifted over the image axes and
    Λ[:, :, i] = U' * (M[i, :] .* U)."""
    M = np.asarray(M)
    d = M.ndim - 1
    Ms = np.fft.ifftshift(M, axes=tuple(range(1, d + 1))).reshape(M.shape[0], -1).astype(np.float64)
    U = np.asarray(U)
    real = not np.iscomplexobj(U)
    U64 = U.astype(np.float64 if real else np.complex128)
    return np.einsum("tq,jt,kt->qjk", Ms, np.conj(U64), U64)


def kmask_of(Lam):
    """kmask = Λ[1, 1, :] .!= 0 in the kernel's element type (ComplexF32 or
    Float32, the eltype of U) and the kept cells (src/FFTNormalOp.jl:44-47),
    0-based."""
    l11 = np.asarray(Lam)[:, 0, 0]
    l11 = l11.astype(np.complex64) if np.iscomplexobj(l11) else l11.astype(np.float32)
    return np.flatnonzero(l11 != 0)


def apply_normal(img_shape, Lam, cmaps, b, alpha=1.0, beta=0.0, x=None):
    """mul!(x, ::_FFTNormalOp, b, α, β) (src/FFTNormalOp.jl:113-149): per coil,
    kL1 = fft(S_c b_k) over the N grid, kL2[q] = Λ[:, :, q] kL1[q] on the
    kmask, x_k += α conj(S_c) ifft(kL2_k).  b, x (K, *N[::-1])."""
    img_shape = tuple(int(v) for v in img_shape)
    d = len(img_shape)
    Lam = np.asarray(Lam)
    K = Lam.shape[1]
    km = kmask_of(Lam)
    b = np.asarray(b, dtype=np.complex128).reshape((K,) + img_shape[::-1])
    out = np.zeros_like(b) if (x is None or beta == 0) else beta * np.asarray(x, dtype=np.complex128)
    maps = [None] if cmaps is None else list(np.asarray(cmaps, dtype=np.complex128))
    axes = tuple(range(1, d + 1))
    for cm in maps:
        kl1 = np.fft.fftn(b if cm is None else cm[None] * b, axes=axes).reshape(K, -1)
        kl2 = np.zeros_like(kl1)
        kl2[:, km] = np.einsum("qjk,kq->jq", Lam[km], kl1[:, km])
        img = np.fft.ifftn(kl2.reshape(b.shape), axes=axes)
        out += alpha * (img if cm is None else np.conj(cm)[None] * img)
    return out


def backprojection(data, trj, cmaps_or_shape, U=None, sample_mask=None):
    """calculate_backprojection for an Integer trajectory (src/BackProjection.jl:
    104-156): per coefficient and coil, dataU[k_idx] += data * conj(U[t, k])
    over the kept samples, ifft (normalised), fftshift; coil-combined with
    conj(S_c) when cmaps are given, else (C, K, *N[::-1])."""
    data = np.asarray(data)
    C, nt, ns = data.shape
    trj = np.asarray(trj)
    if U is None:
        U = np.eye(nt)
    U = np.asarray(U)
    K = U.shape[0]
    per_coil = isinstance(cmaps_or_shape, (tuple, list)) and np.isscalar(cmaps_or_shape[0])
    if per_coil:
        img_shape = tuple(int(s) for s in cmaps_or_shape)
    else:
        cmaps = np.asarray(cmaps_or_shape, dtype=np.complex128)
        img_shape = cmaps.shape[1:][::-1]
    d = len(img_shape)
    G = int(np.prod(img_shape))
    q = sample_cells(trj, img_shape)
    keep = np.ones((nt, ns), bool) if sample_mask is None else np.asarray(sample_mask, bool).reshape(nt, ns)
    axes = tuple(range(1, d + 1))
    imgs = np.zeros((C, K) + img_shape[::-1], dtype=np.complex128)
    for c in range(C):
        grid = np.zeros((K, G), dtype=np.complex128)
        for t in range(nt):
            v = data[c, t][keep[t]].astype(np.complex128)
            for k in range(K):
                np.add.at(grid[k], q[t][keep[t]], v * np.conj(U[k, t]))
        img = np.fft.ifftn(grid.reshape((K,) + img_shape[::-1]), axes=axes)
        imgs[c] = np.fft.fftshift(img, axes=axes)
    if per_coil:
        return imgs
    return np.sum(np.conj(cmaps)[:, None] * imgs, axis=0)
