"""Pins the FP64 oracle against closed forms, brute force and invariants.

The reference ships no golden vectors (every reference test synthesises its
inputs with Julia's RNG), so the oracle is pinned by mathematics chosen so that
a dropped term, a wrong sign or index, or a transposed operand fails a check.
"""
import math

import numpy as np
import pytest

from conftest import load_package, rel_err
from oracle import cg as ocg
from oracle import nufft, toeplitz

synth = load_package().synth


def small_case(seed=0, N=(12, 10), K=3, C=3, nt=6, nspk=2, complex_u=True, d=2):
    rng = np.random.default_rng(seed)
    if d == 2:
        trj = synth.traj_2d_radial_goldenratio(2 * max(N), nspk, nt)
    else:
        trj = synth.traj_kooshball_goldenratio(2 * max(N), nspk, nt).astype(np.float32)
    U = rng.normal(size=(nt, K)) + (1j * rng.normal(size=(nt, K)) if complex_u else 0)
    U, _, _ = np.linalg.svd(U, full_matrices=False)
    U = np.ascontiguousarray(U.T).astype(np.complex64 if complex_u else np.float32)
    cm = synth.gaussian_coil_maps(N, C, seed=seed)
    x = (rng.normal(size=(K,) + N[::-1]) + 1j * rng.normal(size=(K,) + N[::-1]))
    return trj, U, cm, x


def test_nudft_direct_sum():
    rng = np.random.default_rng(1)
    kc = rng.uniform(-0.5, 0.5, (40, 2))
    v = rng.normal(size=(1, 40)) + 1j * rng.normal(size=(1, 40))
    out = nufft.nudft_type1(kc, v, (6, 5), shift=True)
    for iy, ix in [(0, 0), (2, 4), (4, 1)]:
        r = np.array([ix - 3, iy - 2])
        ref = np.sum(v[0] * np.exp(2j * np.pi * kc @ r))
        assert abs(out[0, iy, ix] - ref) < 1e-12 * np.abs(v).sum()
    # FFT-order frequencies
    out = nufft.nudft_type1(kc, v, (6, 5), shift=False)
    r = np.array([-3, -1])  # x index 3 of 6 -> -3; y index 4 of 5 -> -1
    ref = np.sum(v[0] * np.exp(2j * np.pi * kc @ r))
    assert abs(out[0, 4, 3] - ref) < 1e-12 * np.abs(v).sum()


def test_nudft_cartesian_is_dft():
    # k on the Cartesian grid m/N: type-2 with shift equals an fftshifted DFT
    N = 8
    m = (np.arange(N) - N // 2) / N
    kx, ky = np.meshgrid(m, m, indexing="xy")
    kc = np.stack([kx.reshape(-1), ky.reshape(-1)], axis=1)
    rng = np.random.default_rng(2)
    x = rng.normal(size=(1, N, N)) + 1j * rng.normal(size=(1, N, N))
    y = nufft.nudft_type2(kc, x, shift=True).reshape(N, N)
    ref = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(x[0])))
    assert rel_err(y, ref) < 1e-12
    # adjoint then forward: N^2 identity
    xb = nufft.nudft_type1(kc, y.reshape(1, -1), (N, N), shift=True)[0]
    assert rel_err(xb, N * N * x[0]) < 1e-12


@pytest.mark.parametrize("d,n", [(2, (14, 10)), (3, (8, 6, 10))])
def test_nufft64_matches_nudft(d, n):
    rng = np.random.default_rng(3)
    M = 700
    kc = rng.uniform(-0.5, 0.5, (M, d))
    v = rng.normal(size=(2, M)) + 1j * rng.normal(size=(2, M))
    for shift in (True, False):
        assert rel_err(nufft.nufft64_type1(kc, v, n, shift), nufft.nudft_type1(kc, v, n, shift)) < 1e-11
        x = rng.normal(size=(2,) + n[::-1]) + 1j * rng.normal(size=(2,) + n[::-1])
        assert rel_err(nufft.nufft64_type2(kc, x, shift), nufft.nudft_type2(kc, x, shift)) < 1e-11


def test_type1_is_adjoint_of_type2():
    rng = np.random.default_rng(4)
    kc = rng.uniform(-0.5, 0.5, (300, 2))
    x = rng.normal(size=(1, 10, 12)) + 1j * rng.normal(size=(1, 10, 12))
    y = rng.normal(size=(1, 300)) + 1j * rng.normal(size=(1, 300))
    lhs = np.vdot(nufft.nudft_type2(kc, x, True), y)
    rhs = np.vdot(x, nufft.nudft_type1(kc, y, (12, 10), True))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_kmask_contains_sample_cells_and_bounds():
    trj, U, cm, x = small_case()
    kc, _, _ = toeplitz.select_samples(trj)
    n = (24, 20)
    km = toeplitz.kmask_indcs(n, kc)
    assert np.all(np.diff(km) > 0)
    assert km.min() >= 0 and km.max() < n[0] * n[1]
    cells = toeplitz.kmask_cells(kc, n)
    lin = cells[:, 0] + n[0] * cells[:, 1]
    assert np.all(np.isin(lin, km))
    # every cell of the mask lies within the 8-cell box of some sample
    iy, ix = np.divmod(km, n[0])
    dx = np.mod(ix[:, None] - cells[None, :, 0] + 3, n[0])
    dy = np.mod(iy[:, None] - cells[None, :, 1] + 3, n[1])
    assert np.all(np.any((dx < 8) & (dy < 8), axis=1))


def test_kmask_point_symmetric_for_symmetric_trajectory():
    # radial full spokes are point-symmetric; the footprint (base-3..base+4,
    # base = floor(mod(-k,1) n)) of a symmetric point set is symmetric too
    trj = synth.traj_2d_radial_goldenratio(16, 3, 4)
    kc, _, _ = toeplitz.select_samples(trj)
    n = (16, 16)
    km = toeplitz.kmask_indcs(n, kc)
    m = np.zeros(n[::-1], dtype=bool)
    m.reshape(-1)[km] = True
    mirrored = np.roll(np.flip(m, axis=(0, 1)), 1, axis=(0, 1))  # q -> -q mod n
    assert np.array_equal(m, mirrored)


def test_kmask_rejects_non_smooth_grid():
    kc = np.zeros((3, 2), dtype=np.float32)
    with pytest.raises(AssertionError):
        toeplitz.kmask_indcs((14, 16), kc)


def _brute_normal(x, trj, U, cm, mask=None):
    y = toeplitz.forward(x, trj, U, cm, sample_mask=mask)
    return toeplitz.backprojection(y, trj, cm, U=U, sample_mask=mask)


@pytest.mark.parametrize("complex_u", [True, False])
def test_apply_equals_brute_force_without_mask(complex_u):
    N = (12, 10)
    trj, U, cm, x = small_case(N=N, complex_u=complex_u)
    n = (24, 20)
    full = np.arange(n[0] * n[1])
    Lp, km = toeplitz.toeplitz_kernel(N, trj, U, kmask=full)
    got = toeplitz.apply_normal(N, Lp, km, cm, x)
    ref = _brute_normal(x, trj, U, cm)
    assert rel_err(got, ref) < 1e-11


def test_apply_equals_brute_force_3d_and_masked_samples():
    N = (6, 4, 6)
    trj, U, cm, x = small_case(N=N, K=2, C=2, nt=4, nspk=2, d=3)
    mask = np.ones(trj.shape[:2], dtype=bool)
    mask[1, 3:9] = False
    n = tuple(2 * v for v in N)
    full = np.arange(int(np.prod(n)))
    Lp, km = toeplitz.toeplitz_kernel(N, trj, U, sample_mask=mask, kmask=full)
    got = toeplitz.apply_normal(N, Lp, km, cm, x)
    ref = _brute_normal(x, trj, U, cm, mask)
    assert rel_err(got, ref) < 1e-11


def test_full_cartesian_identity():
    # all k = m/N, one frame, K = 1, U = 1, RSS-normalised coils -> prod(N) I
    N = 8
    m = (np.arange(N) - N // 2) / N
    kx, ky = np.meshgrid(m, m, indexing="xy")
    trj = np.stack([kx.reshape(-1), ky.reshape(-1)], axis=1)[None].astype(np.float32)
    U = np.ones((1, 1), dtype=np.complex64)
    cm = synth.gaussian_coil_maps((N, N), 4)
    Lp, km = toeplitz.toeplitz_kernel((N, N), trj, U)
    assert km.size == (2 * N) ** 2  # the footprint covers the whole grid
    rng = np.random.default_rng(5)
    x = rng.normal(size=(1, N, N)) + 1j * rng.normal(size=(1, N, N))
    got = toeplitz.apply_normal((N, N), Lp, km, cm, x)
    assert rel_err(got, N * N * x) < 1e-6  # cmaps are float32-rounded RSS-1 maps


def test_kernel_hermitian_and_real_invariants():
    N = (10, 10)
    trj, U, cm, x = small_case(N=N, K=3)
    Lp, km = toeplitz.toeplitz_kernel(N, trj, U)
    full = toeplitz.unpack(Lp, 3)
    herm = np.conj(np.transpose(full, (0, 2, 1)))
    off = ~np.eye(3, dtype=bool)
    assert np.allclose(full[:, off], herm[:, off])
    # the diagonal PSF p_jj is Hermitian except on its Nyquist planes (index -N),
    # so Im(diag) equals the imaginary part of the transform of those planes
    kc, frame, _ = toeplitz.select_samples(trj)
    n = (20, 20)
    for j in range(3):
        p = toeplitz.psf_grid(kc, frame, U, n, j, j)
        pn = np.zeros_like(p)
        pn[10, :] = p[10, :]
        pn[:, 10] = p[:, 10]
        lam_n = np.fft.fft2(pn).reshape(-1)[km]
        assert np.allclose(full[:, j, j].imag, lam_n.imag, atol=1e-9 * np.abs(full).max())
    trj, Ur, cm, x = small_case(N=N, K=3, complex_u=False)
    Lr, _ = toeplitz.toeplitz_kernel(N, trj, Ur)
    assert not np.iscomplexobj(Lr)
    fr = toeplitz.unpack(Lr, 3)
    assert np.allclose(fr, np.transpose(fr, (0, 2, 1)))


def test_k1_reduces_to_classic_toeplitz():
    # K = 1, U = 1: the kernel is FFT of the plain PSF sum_j exp(2 pi i k_j r)
    N = (8, 8)
    trj = synth.traj_2d_radial_goldenratio(16, 4, 1)
    U = np.ones((1, 1), dtype=np.complex64)
    kc, _, _ = toeplitz.select_samples(trj)
    n = (16, 16)
    full = np.arange(256)
    Lp, km = toeplitz.toeplitz_kernel(N, trj, U, kmask=full)
    r = np.fft.fftfreq(16, 1 / 16)
    ry, rx = np.meshgrid(r, r, indexing="ij")
    p = np.exp(2j * np.pi * (kc[:, 0, None, None] * rx[None] + kc[:, 1, None, None] * ry[None])).sum(0)
    assert rel_err(Lp[:, 0], np.fft.fft2(p).reshape(-1)) < 1e-12


def test_apply_adjoint_psd_alpha_beta():
    N = (10, 8)
    trj, U, cm, x = small_case(N=N)
    Lp, km = toeplitz.toeplitz_kernel(N, trj, U)
    rng = np.random.default_rng(6)
    y = rng.normal(size=x.shape) + 1j * rng.normal(size=x.shape)
    Ax = toeplitz.apply_normal(N, Lp, km, cm, x)
    Ay = toeplitz.apply_normal(N, Lp, km, cm, y)
    assert abs(np.vdot(y, Ax) - np.vdot(Ay, x)) < 1e-10 * abs(np.vdot(y, Ax))
    assert np.vdot(x, Ax).real > 0
    z = rng.normal(size=x.shape) + 0j
    got = toeplitz.apply_normal(N, Lp, km, cm, x, alpha=0.5 - 2j, beta=1.5, x=z)
    assert rel_err(got, 1.5 * z + (0.5 - 2j) * Ax) < 1e-13
    # beta == 0 overwrites NaN
    got = toeplitz.apply_normal(N, Lp, km, cm, x, beta=0.0, x=np.full_like(z, np.nan))
    assert rel_err(got, Ax) < 1e-13


def test_backprojection_brute_force_and_dcf():
    N = (8, 6)
    trj, U, cm, x = small_case(N=N, K=2, C=2, nt=3)
    rng = np.random.default_rng(7)
    data = rng.normal(size=(2,) + trj.shape[:2]) + 1j * rng.normal(size=(2,) + trj.shape[:2])
    for dc in ["none", "radial_2D", "radial_3D"]:
        got = toeplitz.backprojection(data, trj, cm, U=U, density_compensation=dc)
        ref = np.zeros((2,) + N[::-1], dtype=complex)
        for it in range(trj.shape[0]):
            for js in range(trj.shape[1]):
                k = trj[it, js].astype(np.float64)
                w = {"none": 1.0, "radial_2D": np.linalg.norm(k), "radial_3D": k @ k}[dc]
                for iy in range(N[1]):
                    for ix in range(N[0]):
                        e = np.exp(2j * np.pi * (k[0] * (ix - N[0] // 2) + k[1] * (iy - N[1] // 2)))
                        for kk in range(2):
                            ref[kk, iy, ix] += np.sum(np.conj(cm[:, iy, ix]) * data[:, it, js]) \
                                * w * np.conj(U[kk, it]) * e
        assert rel_err(got, ref) < 1e-11
    with pytest.raises(ValueError):
        toeplitz.backprojection(data, trj, cm, U=U, density_compensation="bogus")
    # per-coil variant keeps the coil axis and drops cmaps
    pc = toeplitz.backprojection(data, trj, N, U=U)
    assert pc.shape == (2, 2) + N[::-1]
    assert rel_err(np.sum(np.conj(cm)[:, None] * pc, axis=0),
                   toeplitz.backprojection(data, trj, cm, U=U)) < 1e-12


def test_cartesian_backprojection_recovers_image():
    # fully sampled Cartesian data backprojects to prod(N) x (test/backprojection_cart.jl)
    N = 8
    m = (np.arange(N) - N // 2) / N
    kx, ky = np.meshgrid(m, m, indexing="xy")
    trj = np.stack([kx.reshape(-1), ky.reshape(-1)], axis=1)[None].astype(np.float32)
    cm = synth.gaussian_coil_maps((N, N), 3).astype(np.complex128)
    rng = np.random.default_rng(8)
    x = rng.normal(size=(1, N, N)) + 0j
    U = np.ones((1, 1))
    data = toeplitz.forward(x, trj, U, cm)
    b = toeplitz.backprojection(data, trj, cm, U=U)
    assert rel_err(b / N ** 2, x) < 1e-6  # cmaps are float32-rounded RSS-1 maps


def test_cg_matches_textbook_and_solve():
    rng = np.random.default_rng(9)
    n = 30
    B = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    A = B.conj().T @ B + n * np.eye(n)
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    hist = []
    x = ocg.cg(lambda v: A @ v, b, maxiter=200, reltol=0.0, history=hist)
    assert rel_err(x, np.linalg.solve(A, b)) < 1e-10
    # textbook (Hestenes-Stiefel) CG for 5 iterations gives the same iterate
    xt = np.zeros(n, complex)
    r = b.copy()
    p = r.copy()
    for _ in range(5):
        Ap = A @ p
        a = np.vdot(r, r) / np.vdot(p, Ap)
        xt = xt + a * p
        rn = r - a * Ap
        p = rn + (np.vdot(rn, rn) / np.vdot(r, r)) * p
        r = rn
    x5 = ocg.cg(lambda v: A @ v, b, maxiter=5, reltol=0.0)
    assert rel_err(x5, xt) < 1e-12
    # early stop at reltol: the residual when stopping is below tol
    hist = []
    ocg.cg(lambda v: A @ v, b, maxiter=1000, history=hist)
    assert hist[-1] <= ocg.REL_TOL_F32 * np.linalg.norm(b)
    assert len(hist) < 1000


def test_recon_recovers_truth_like_reference_test():
    # test/reconstruct_radial.jl:77-96 restated: CG(20) on simulated data
    # approaches the k-space-disk-cropped truth
    N = 16
    nt, ncyc, K = 10, 8, 2
    trj = synth.traj_2d_radial_goldenratio(2 * N, ncyc, nt)
    U = synth.t2_basis(nt, K)
    cm = synth.gaussian_coil_maps((N, N), 4)
    x = synth.ellipse_phantom((N, N), K, seed=3).astype(np.complex128)
    data = toeplitz.forward(x, trj, U, cm)
    b = toeplitz.backprojection(data, trj, cm, U=U)
    Lp, km = toeplitz.toeplitz_kernel((N, N), trj, U)
    xr = ocg.cg(lambda v: toeplitz.apply_normal((N, N), Lp, km, cm, v.reshape(x.shape)).reshape(-1),
                b.reshape(-1), maxiter=20).reshape(x.shape)
    xc = np.fft.fftshift(np.fft.fft2(x), axes=(1, 2))
    iy, ix = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    xc[:, (ix - N / 2) ** 2 + (iy - N / 2) ** 2 > (N / 2) ** 2] = 0
    xc = np.fft.ifft2(np.fft.ifftshift(xc, axes=(1, 2)))
    # coefficient 1 carries most energy; compare like the reference (rtol 1e-1)
    assert rel_err(xr, xc) < 0.2


def test_cpu_port_matches_oracle():
    """The FP32 CPU-baseline port computes the same operator and
    backprojection as the FP64 oracle (to FP32 accuracy)."""
    from oracle import cpu_port
    N = (16, 16)
    trj, U, cm, x = small_case(N=N, K=3, C=3, nt=6, nspk=3)
    Lp, km = toeplitz.toeplitz_kernel(N, trj, U)
    op = cpu_port.CpuNormalOp(N, Lp, km, cm, workers=2)
    got = op.apply(x.astype(np.complex64)).reshape(x.shape)
    ref = toeplitz.apply_normal(N, Lp, km, cm.astype(np.complex128), x)
    assert rel_err(got, ref) < 1e-5
    rng = np.random.default_rng(3)
    data = (rng.normal(size=(3,) + trj.shape[:2]) + 1j * rng.normal(size=(3,) + trj.shape[:2])).astype(np.complex64)
    bp = cpu_port.backprojection_port(data, trj, cm, U, workers=2)
    assert rel_err(bp, toeplitz.backprojection(data, trj, cm, U=U)) < 1e-5
