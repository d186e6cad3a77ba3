"""FP64 non-uniform Fourier sums (test infrastructure only).

Sign convention of the reference (NonuniformFFTs.jl with
``_transform_point_convention``, comment at src/BackProjection.jl:35):
  type 1 (adjoint, backprojection):  out[r] = sum_j v_j exp(+2 pi i k_j . r)
  type 2 (forward, data simulation): y_j    = sum_r x[r] exp(-2 pi i k_j . r)
Frequencies r per dim of size n: ``shift=True`` -> r = idx - n//2 (fftshift
ordering, src/BackProjection.jl:75); ``shift=False`` -> FFT order
(0..n/2-1, -n/2..-1), used for the Toeplitz PSF (src/NFFTNormalOp.jl:142).

Arrays: coordinates kc (M, d) with column 0 = x; grids are row-major arrays
of shape n[::-1] (x fastest), i.e. the memory of a Julia array of size n.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.fft as sfft

from . import native


def freqs(n: int, shift: bool) -> np.ndarray:
    if shift:
        return np.arange(n, dtype=np.float64) - (n // 2)
    return np.fft.fftfreq(n, 1.0 / n)


def _chunks(M, size):
    for a in range(0, M, size):
        yield slice(a, min(M, a + size))


def nudft_type1(kc, v, n, shift, chunk=4096):
    """Exact type-1 sum.  kc (M, d); v (nch, M) complex; n (x-first sizes).
    Returns (nch, *n[::-1]) complex128."""
    kc = np.asarray(kc, dtype=np.float64)
    v = np.atleast_2d(np.asarray(v, dtype=np.complex128))
    M, d = kc.shape
    nch = v.shape[0]
    rs = [freqs(n[i], shift) for i in range(d)]
    out = np.zeros((nch,) + tuple(n[::-1]), dtype=np.complex128)
    for sl in _chunks(M, chunk):
        E = [np.exp(2j * math.pi * kc[sl, i, None] * rs[i][None, :]) for i in range(d)]
        if d == 1:
            out += v[:, sl] @ E[0]
        elif d == 2:
            W = v[:, sl, None] * E[1][None]          # (nch, m, ny)
            out += np.swapaxes(W, 1, 2) @ E[0]       # (nch, ny, nx)
        else:
            W = v[:, sl, None, None] * E[2][None, :, :, None] * E[1][None, :, None, :]
            m = W.shape[1]
            W = W.reshape(nch, m, -1)
            out += (np.swapaxes(W, 1, 2) @ E[0]).reshape(out.shape)
    return out


def nudft_type2(kc, x, shift, chunk=4096):
    """Exact type-2 sum.  x (nch, *n[::-1]).  Returns (nch, M) complex128."""
    kc = np.asarray(kc, dtype=np.float64)
    x = np.asarray(x, dtype=np.complex128)
    M, d = kc.shape
    nch = x.shape[0]
    n = x.shape[1:][::-1]
    rs = [freqs(n[i], shift) for i in range(d)]
    out = np.empty((nch, M), dtype=np.complex128)
    for sl in _chunks(M, chunk):
        E = [np.exp(-2j * math.pi * kc[sl, i, None] * rs[i][None, :]) for i in range(d)]
        if d == 1:
            out[:, sl] = x @ E[0].T
        elif d == 2:
            T = x @ E[0].T                            # (nch, ny, m)
            out[:, sl] = np.einsum("cym,my->cm", T, E[1])
        else:
            T = x @ E[0].T                            # (nch, nz, ny, m)
            T = np.einsum("czym,my->czm", T, E[1])
            out[:, sl] = np.einsum("czm,mz->cm", T, E[2])
    return out


# --------------------------------------------------------------------------
# High-accuracy FP64 NUFFT (exponential-of-semicircle kernel, sigma = 2)
# --------------------------------------------------------------------------
W64 = 14
BETA64 = 2.30 * W64


def es_phi(t, w, beta):
    z = 2.0 * np.asarray(t, dtype=np.float64) / w
    a = 1.0 - z * z
    out = np.zeros_like(a)
    m = a > 0
    out[m] = np.exp(beta * (np.sqrt(a[m]) - 1.0))
    return out


def es_phihat(r, n_os, w, beta, nodes=400):
    """Fourier transform of the ES kernel at integer frequency r on an n_os
    grid: int phi(t) exp(2 pi i r t / n_os) dt (real: phi is even), by
    Gauss-Legendre quadrature."""
    x, wq = np.polynomial.legendre.leggauss(nodes)
    t = 0.5 * w * x
    wq = 0.5 * w * wq
    ph = es_phi(t, w, beta)
    r = np.asarray(r, dtype=np.float64)
    return (np.cos(2 * math.pi * np.outer(r, t) / n_os) * (ph * wq)[None]).sum(axis=1)


def nufft64_type1(kc, v, n, shift, w=W64, beta=BETA64):
    kc = np.asarray(kc, dtype=np.float64)
    v = np.atleast_2d(np.asarray(v, dtype=np.complex128))
    M, d = kc.shape
    nch = v.shape[0]
    nos = [2 * n[i] for i in range(d)]
    s = np.mod(kc, 1.0) * np.array(nos, dtype=np.float64)[None, :]
    s = np.where(s >= np.array(nos)[None, :], 0.0, s)
    g = native.es_spread_f64(s, v, nos, w, beta).reshape((nch,) + tuple(nos[::-1]))
    axes = tuple(range(1, d + 1))
    G = sfft.ifftn(g, axes=axes, workers=-1) * float(np.prod(nos))
    out = G
    for i in range(d):
        r = freqs(n[i], shift).astype(np.int64)
        idx = np.mod(r, nos[i])
        ax = d - i  # array axis of dim i
        out = np.take(out, idx, axis=ax)
        ph = es_phihat(r, nos[i], w, beta)
        shape = [1] * (d + 1)
        shape[ax] = n[i]
        out = out / ph.reshape(shape)
    return out


def nufft64_type2(kc, x, shift, w=W64, beta=BETA64):
    kc = np.asarray(kc, dtype=np.float64)
    x = np.asarray(x, dtype=np.complex128)
    M, d = kc.shape
    nch = x.shape[0]
    n = x.shape[1:][::-1]
    nos = [2 * n[i] for i in range(d)]
    xt = x.copy()
    big = np.zeros((nch,) + tuple(nos[::-1]), dtype=np.complex128)
    idxs = []
    for i in range(d):
        r = freqs(n[i], shift).astype(np.int64)
        ax = d - i
        ph = es_phihat(r, nos[i], w, beta)
        shape = [1] * (d + 1)
        shape[ax] = n[i]
        xt = xt / ph.reshape(shape)
        idxs.append(np.mod(r, nos[i]))
    # place into the oversampled grid
    ix = np.ix_(*([np.arange(nch)] + [idxs[d - 1 - a] for a in range(d)]))
    big[ix] = xt
    axes = tuple(range(1, d + 1))
    g = sfft.fftn(big, axes=axes, workers=-1).reshape(nch, -1)
    s = np.mod(kc, 1.0) * np.array(nos, dtype=np.float64)[None, :]
    s = np.where(s >= np.array(nos)[None, :], 0.0, s)
    return native.es_interp_f64(s, g, nos, w, beta)


def type1(kc, v, n, shift, exact_limit=2e9):
    """Exact NUDFT when affordable (M * prod(n) * nch below exact_limit), else
    the FP64 NUFFT."""
    v = np.atleast_2d(v)
    cost = float(np.asarray(kc).shape[0]) * float(np.prod(n)) * v.shape[0]
    if cost <= exact_limit:
        return nudft_type1(kc, v, n, shift)
    return nufft64_type1(kc, v, n, shift)


def type2(kc, x, shift, exact_limit=2e9):
    x = np.asarray(x)
    cost = float(np.asarray(kc).shape[0]) * float(np.prod(x.shape[1:])) * x.shape[0]
    if cost <= exact_limit:
        return nudft_type2(kc, x, shift)
    return nufft64_type2(kc, x, shift)
