"""FP32 CPU port of the reference's CPU path, timed as the CPU baseline.

TEST INFRASTRUCTURE ONLY (imported by the benchmark's CPU-baseline leg).

The reference CPU path cannot run here (Julia is absent), so this restates its
algorithm and cost structure in ComplexF32 with NumPy/SciPy:
  * apply: src/NFFTNormalOp.jl:227-267 -- per coil, per coefficient zero-pad
    S_c b_k into the 2N grid and FFT (threaded FFTs), per-voxel K x K multiply
    over the kmask (full (K, K, Nkm) kernel) split over the host threads as
    the reference's @tasks loop splits it (:254), inverse FFT per
    coefficient, crop and accumulate conj(S_c) times the result.
  * backprojection: src/BackProjection.jl:62-102 -- one type-1 NUFFT per
    (coefficient, coil): data * conj(U[t, k]), spread onto the sigma = 2 grid
    with an 8-point kernel (NonuniformFFTs.jl half-support 4), FFT, crop and
    deapodise, then accumulate conj(S_c) times the image.
  * CG: the IterativeSolvers.jl recurrence (oracle/cg.py) with ComplexF32
    vectors.
"""
from __future__ import annotations

import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import scipy.fft as sfft

from . import native, nufft, toeplitz

W_PORT = 8
BETA_PORT = 2.30 * W_PORT


class CpuNormalOp:
    def __init__(self, img_shape, Lp, kmask, cmaps, workers=None):
        self.N = tuple(int(v) for v in img_shape)
        self.n = tuple(2 * v for v in self.N)
        self.K = (math.isqrt(8 * Lp.shape[1] + 1) - 1) // 2
        self.full = toeplitz.unpack(np.asarray(Lp), self.K).astype(np.complex64)  # (Nkm, K, K) [i, col, row]
        self.kmask = np.asarray(kmask)
        self.cmaps = np.asarray(cmaps, dtype=np.complex64)
        self.workers = workers or os.cpu_count()
        # voxel chunks of the K x K multiply, one task each (numpy's matmul
        # loops release the GIL, so the threads run in parallel)
        nkm = self.kmask.size
        edges = np.linspace(0, nkm, self.workers + 1).astype(np.int64)
        self.chunks = [slice(int(a), int(b)) for a, b in zip(edges[:-1], edges[1:]) if b > a]
        self.pool = ThreadPoolExecutor(max_workers=self.workers)

    def _multiply(self, v):
        """o[i, j] = sum_k Lam_i[j, k] v[i, k] over the kmask voxels, in
        parallel voxel chunks."""
        o = np.empty_like(v)

        def run(sl):
            o[sl] = np.matmul(v[sl, None, :], self.full[sl])[:, 0, :]

        list(self.pool.map(run, self.chunks))
        return o

    def apply(self, b):
        K = self.K
        b = b.reshape((K,) + self.N[::-1]).astype(np.complex64)
        x = np.zeros_like(b)
        crop = (slice(None),) + tuple(slice(0, v) for v in self.N[::-1])
        axes = tuple(range(1, len(self.N) + 1))
        G = int(np.prod(self.n))
        for cm in self.cmaps:
            kl1 = np.zeros((K,) + self.n[::-1], dtype=np.complex64)
            kl1[crop] = cm[None] * b
            kl1 = sfft.fftn(kl1, axes=axes, workers=self.workers)
            v = kl1.reshape(K, G)[:, self.kmask].T                       # (Nkm, K)
            o = self._multiply(v)                                        # row j = sum_k Lam[j,k] v_k
            kl2 = np.zeros((K, G), dtype=np.complex64)
            kl2[:, self.kmask] = o.T
            img = sfft.ifftn(kl2.reshape((K,) + self.n[::-1]), axes=axes, workers=self.workers)
            x += np.conj(cm)[None] * img[crop]
        return x.reshape(-1)


def backprojection_port(data, trj, cmaps, U, workers=None):
    """K*C single-channel type-1 NUFFTs in FP32 (src/BackProjection.jl:62-102)."""
    workers = workers or os.cpu_count()
    kc, frame, _ = toeplitz.select_samples(trj)
    cmaps = np.asarray(cmaps, dtype=np.complex64)
    N = cmaps.shape[1:][::-1]
    d = len(N)
    nos = [2 * v for v in N]
    s = (np.mod(kc.astype(np.float64), 1.0) * np.array(nos)[None]).astype(np.float32)
    C = data.shape[0]
    K = U.shape[0]
    yr = data.reshape(C, -1)
    scratch = np.empty(workers * 2 * int(np.prod(nos)), dtype=np.float32)
    ph = [nufft.es_phihat(nufft.freqs(N[i], True), nos[i], W_PORT, BETA_PORT) for i in range(d)]
    deap = np.ones(N[::-1])
    for i in range(d):
        shape = [1] * d
        shape[d - 1 - i] = N[i]
        deap = deap / ph[i].reshape(shape)
    idx = [np.mod(nufft.freqs(N[i], True).astype(np.int64), nos[i]) for i in range(d)]
    ix = np.ix_(*[idx[d - 1 - a] for a in range(d)])
    out = np.zeros((K,) + tuple(N[::-1]), dtype=np.complex64)
    Uc = np.conj(U[:, frame]).astype(np.complex64)
    for k in range(K):
        for c in range(C):
            v = (yr[c] * Uc[k]).astype(np.complex64)
            g = native.spread_f32_1ch(s, v, nos, W_PORT, BETA_PORT, scratch, workers)
            G = sfft.ifftn(g.reshape(nos[::-1]), workers=workers) * np.float32(np.prod(nos))
            img = (G[ix] * deap).astype(np.complex64)
            out[k] += np.conj(cmaps[c]) * img
    return out
