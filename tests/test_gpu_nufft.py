"""GPU parity of the backprojection, the on-device kmask / Toeplitz-kernel
construction and the end-to-end reconstruction against the FP64 oracle.

Tolerances (relative L2): kmask bit-exact (integer indices); backprojection and
kernel 1e-5 (FP32 type-1 NUFFT with a width-8 ES kernel: ~1.2e-7 kernel error,
~4e-7 measured with FP32 rounding); a CG reconstruction 1e-4 against the FP64
pipeline on the same inputs."""
import numpy as np
import pytest

from conftest import load_package, rel_err
from oracle import cg as ocg
from oracle import toeplitz

pytestmark = pytest.mark.gpu

NUFFT_TOL = 1e-5
# the kernel is built in FP64 and rounded once to ComplexF32: what remains is
# the W = 8 spreading kernel's truncation error (~1e-7) and the FP32 rounding
KERNEL_TOL = 3e-7
RECON_TOL = 1e-4


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.fixture(scope="module")
def msr():
    return load_package()


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def radial_case(N, K, C, nt, nspk, seed=0, real=False, half=False):
    synth = load_package().synth
    trj = synth.traj_2d_radial_goldenratio(2 * max(N), nspk, nt)
    if half:  # asymmetric half-spokes (test/reconstruct_radial_asym.jl:46-55)
        trj = trj.reshape(nt, nspk, 2 * max(N), 2)[:, :, :max(N)].reshape(nt, -1, 2).copy()
    U = synth.t2_basis(nt, K, seed=seed, complex_out=not real)
    cm = synth.gaussian_coil_maps(N, C, seed=seed)
    x = synth.ellipse_phantom(N, K, seed=seed)
    return trj, U, cm, x


KERNEL_CASES = [
    ((16, 16), 3, 8, 4, False, False, False),
    ((16, 16), 3, 8, 4, True, False, False),
    ((32, 16), 2, 6, 3, False, True, False),
    ((24, 40), 4, 5, 4, False, False, True),
    ((48, 48), 5, 10, 2, True, False, False),
]


@pytest.mark.parametrize("N,K,nt,nspk,real,half,masked", KERNEL_CASES)
def test_kernel_and_kmask_match_oracle(torch, msr, N, K, nt, nspk, real, half, masked):
    print(f"kernel case N={N} K={K} real={real} half={half} masked={masked}", flush=True)
    trj, U, cm, x = radial_case(N, K, 2, nt, nspk, real=real, half=half)
    mask = None
    if masked:
        mask = np.ones(trj.shape[:2], dtype=bool)
        mask[0, 5:40] = False  # drop part of a spoke (test/data_removal.jl)
    A = msr.NFFTNormalOp(N, dev(torch, trj), dev(torch, U), cmaps=[dev(torch, c) for c in cm],
                         sample_mask=None if mask is None else dev(torch, mask))
    lam_ref, km_ref = toeplitz.toeplitz_kernel(N, trj, U, sample_mask=mask)
    km = A.prod.A.kmask_indcs.cpu().numpy()
    assert np.array_equal(km, km_ref + 1)
    lam = A.prod.A.Λ.cpu().numpy()
    assert lam.dtype == (np.float32 if real else np.complex64)
    print(f"kernel rel err {rel_err(lam, lam_ref):.3e}", flush=True)
    assert rel_err(lam, lam_ref) < KERNEL_TOL


BP_CASES = [
    ((16, 16), 3, 4, 8, 4, "none", False),
    ((32, 16), 2, 3, 6, 3, "radial_2D", True),
    ((24, 40), 4, 2, 5, 4, "radial_3D", False),
    ((48, 48), 5, 6, 10, 2, "none", False),
    # 64 x 64 grid, 16k samples: the four centre tiles hold more than 4096
    # (point, tile) pairs and are split into work items summed afterwards;
    # 30 and 60 channels make two groups of 16 and three of 25 (padded)
    ((32, 32), 5, 6, 64, 4, "radial_2D", False),
    ((32, 32), 5, 12, 64, 4, "none", True),
]


@pytest.mark.parametrize("N,K,C,nt,nspk,dcf,masked", BP_CASES)
def test_backprojection_matches_oracle(torch, msr, N, K, C, nt, nspk, dcf, masked):
    print(f"bp case N={N} K={K} C={C} dcf={dcf} masked={masked}", flush=True)
    trj, U, cm, x = radial_case(N, K, C, nt, nspk)
    rng = np.random.default_rng(1)
    data = (rng.normal(size=(C,) + trj.shape[:2]) + 1j * rng.normal(size=(C,) + trj.shape[:2])).astype(np.complex64)
    mask = None
    if masked:
        mask = np.ones(trj.shape[:2], dtype=bool)
        mask[1, ::3] = False
    args = dict(U=dev(torch, U), density_compensation=dcf,
                sample_mask=None if mask is None else dev(torch, mask))
    b = msr.calculate_backprojection(dev(torch, data), dev(torch, trj), [dev(torch, c) for c in cm],
                                     **args).cpu().numpy()
    ref = toeplitz.backprojection(data, trj, cm, U=U, sample_mask=mask, density_compensation=dcf)
    print(f"bp rel err {rel_err(b, ref):.3e}", flush=True)
    assert rel_err(b, ref) < NUFFT_TOL
    pc = msr.calculate_backprojection(dev(torch, data), dev(torch, trj), N, **args).cpu().numpy()
    ref_pc = toeplitz.backprojection(data, trj, N, U=U, sample_mask=mask, density_compensation=dcf)
    assert pc.shape == (C, K) + N[::-1]
    assert rel_err(pc, ref_pc) < NUFFT_TOL


@pytest.mark.parametrize("real", [False, True])
def test_channelwise_backprojection_matches_fused(torch, msr, real):
    """The reference's call sequence (one single-channel NUFFT per (k, c))
    gives the fused multi-channel result, and both match the FP64 oracle."""
    N, K, C = (32, 32), 4, 5
    trj, U, cm, x = radial_case(N, K, C, 64, 4, real=real)
    rng = np.random.default_rng(3)
    data = (rng.normal(size=(C,) + trj.shape[:2]) + 1j * rng.normal(size=(C,) + trj.shape[:2])).astype(np.complex64)
    args = dict(U=dev(torch, U), density_compensation="radial_2D")
    cml = [dev(torch, c) for c in cm]
    fused = msr.calculate_backprojection(dev(torch, data), dev(torch, trj), cml, **args).cpu().numpy()
    cw = msr.calculate_backprojection(dev(torch, data), dev(torch, trj), cml, channelwise=True,
                                      **args).cpu().numpy()
    ref = toeplitz.backprojection(data, trj, cm, U=U, density_compensation="radial_2D")
    print(f"channelwise vs fused {rel_err(cw, fused):.3e}, vs oracle {rel_err(cw, ref):.3e}", flush=True)
    assert rel_err(cw, fused) < 1e-6
    assert rel_err(cw, ref) < NUFFT_TOL
    with pytest.raises(msr.ArgumentError):
        msr.calculate_backprojection(dev(torch, data), dev(torch, trj), cml, channelwise=True)


def test_backprojection_identity_basis_and_array_dcf(torch, msr):
    N = (16, 16)
    trj, U, cm, x = radial_case(N, 2, 2, 3, 2)
    rng = np.random.default_rng(2)
    data = (rng.normal(size=(2,) + trj.shape[:2]) + 1j * rng.normal(size=(2,) + trj.shape[:2])).astype(np.complex64)
    w = rng.uniform(0.5, 2.0, size=trj.shape[0] * trj.shape[1]).astype(np.float32)
    b = msr.calculate_backprojection(dev(torch, data), dev(torch, trj), [dev(torch, c) for c in cm],
                                     density_compensation=dev(torch, w)).cpu().numpy()
    ref = toeplitz.backprojection(data, trj, cm, U=None, density_compensation=w)
    assert b.shape == (3,) + N[::-1]  # U = I(Nt): one coefficient per frame
    assert rel_err(b, ref) < NUFFT_TOL


def test_backprojection_errors(torch, msr):
    N = (16, 16)
    trj, U, cm, x = radial_case(N, 2, 2, 3, 2)
    data = np.zeros((2,) + trj.shape[:2], np.complex64)
    with pytest.raises(msr.ErrorException):
        msr.calculate_backprojection(dev(torch, data), dev(torch, trj), [dev(torch, c) for c in cm],
                                     U=dev(torch, U), density_compensation="bogus")
    with pytest.raises(msr.ArgumentError):
        msr.calculate_backprojection(dev(torch, data[:1]), dev(torch, trj), [dev(torch, c) for c in cm],
                                     U=dev(torch, U))
    with pytest.raises(msr.ArgumentError):
        msr.calculate_backprojection(dev(torch, data), dev(torch, trj), [dev(torch, c) for c in cm],
                                     U=dev(torch, U[:, :2]))
    with pytest.raises(AssertionError):
        msr.NFFTNormalOp(N, dev(torch, trj), dev(torch, U),
                         sample_mask=dev(torch, np.zeros(trj.shape[:2], bool)))


def config1_case(real=False, seed=1001):
    """Benchmark config 1: 64x64, 8 Gaussian coils, golden-angle radial with 128
    samples/spoke, 64 frames x 8 spokes, K = 4 T2 basis, ellipse coefficient
    maps, data = forward model + complex noise at 30 dB."""
    synth = load_package().synth
    N = (64, 64)
    trj = synth.traj_2d_radial_goldenratio(128, 8, 64)
    U = synth.t2_basis(64, 4, seed=seed, complex_out=not real)
    cm = synth.gaussian_coil_maps(N, 8, seed=seed)
    x = synth.ellipse_phantom(N, 4, seed=seed)
    data = toeplitz.forward(x.astype(np.complex128), trj, U, cm.astype(np.complex128))
    rng = np.random.default_rng(seed)
    sig = np.sqrt(np.mean(np.abs(data) ** 2)) * 10 ** (-30 / 20) / np.sqrt(2)
    data = (data + sig * (rng.normal(size=data.shape) + 1j * rng.normal(size=data.shape))).astype(np.complex64)
    return N, trj, U, cm, data


@pytest.mark.parametrize("real", [False, True])
def test_recon_config1_matches_fp64(torch, msr, real):
    """Config 1 end to end on the device (BP, kernel build, CG(20)) against
    (a) an FP64 CG on the same ComplexF32 operator and right-hand side the
    device built (solve parity) and (b) the whole pipeline in FP64 from the
    trajectory (end-to-end parity)."""
    print(f"config1 recon real={real}", flush=True)
    N, trj, U, cm, data = config1_case(real)
    cml = [dev(torch, c) for c in cm]
    b = msr.calculate_backprojection(dev(torch, data), dev(torch, trj), cml, U=dev(torch, U))
    A = msr.NFFTNormalOp(N, dev(torch, trj), dev(torch, U), cmaps=cml)
    xr = msr.cg(A, b.reshape(-1), maxiter=20, reltol=0.0).cpu().numpy()
    lam_d = A.prod.A.Λ.cpu().numpy()
    km_d = A.prod.A.kmask_indcs.cpu().numpy() - 1
    bd = b.cpu().numpy().astype(np.complex128)
    cm2 = cm.astype(np.complex128)
    Ad = lambda v: toeplitz.apply_normal(N, lam_d.astype(np.float64 if real else np.complex128),
                                         km_d, cm2, v.reshape(bd.shape)).reshape(-1)
    ref_solve = ocg.cg(Ad, bd.reshape(-1), maxiter=20, reltol=0.0)
    b64 = toeplitz.backprojection(data, trj, cm, U=U)
    lam, km = toeplitz.toeplitz_kernel(N, trj, U)
    Af = lambda v: toeplitz.apply_normal(N, lam, km, cm2, v.reshape(b64.shape)).reshape(-1)
    ref = ocg.cg(Af, b64.reshape(-1), maxiter=20, reltol=0.0)
    xr32 = msr.cg(A, b.reshape(-1), maxiter=20, reltol=0.0, precision="f32").cpu().numpy()
    print(f"config1 rel err: solve {rel_err(xr, ref_solve):.3e}, end-to-end {rel_err(xr, ref):.3e} "
          f"(fp32 CG vectors {rel_err(xr32, ref):.3e}); bp {rel_err(bd, b64):.3e}, "
          f"kernel {rel_err(lam_d, lam):.3e}", flush=True)
    assert np.array_equal(km_d, km)
    assert rel_err(xr, ref_solve) < RECON_TOL
    assert rel_err(xr, ref) < RECON_TOL


def test_recon_undersampled_small_case(torch, msr):
    """A small, strongly undersampled case (32x32, 36 spokes of 64 samples,
    K = 3): CG(20) amplifies FP32-level differences ~100-500x here (measured on
    the CPU: 1e-7 relative noise in the kernel moves the iterate by 1.3e-4).
    With the FP64 apply (the default precision) the solve on the device-built
    operator is held to 1e-7 and the end-to-end result (FP32 kernel and
    backprojection) to 1e-4; the FP32-apply precisions are reported."""
    N = (32, 32)
    K, C, nt, nspk = 3, 4, 12, 3
    trj, U, cm, x = radial_case(N, K, C, nt, nspk)
    data = toeplitz.forward(x.astype(np.complex128), trj, U, cm.astype(np.complex128)).astype(np.complex64)
    cml = [dev(torch, c) for c in cm]
    b = msr.calculate_backprojection(dev(torch, data), dev(torch, trj), cml, U=dev(torch, U))
    A = msr.NFFTNormalOp(N, dev(torch, trj), dev(torch, U), cmaps=cml)
    xr = msr.cg(A, b.reshape(-1), maxiter=20, reltol=0.0).cpu().numpy()
    lam_d = A.prod.A.Λ.cpu().numpy().astype(np.complex128)
    km_d = A.prod.A.kmask_indcs.cpu().numpy() - 1
    bd = b.cpu().numpy().astype(np.complex128)
    cm2 = cm.astype(np.complex128)
    Ad = lambda v: toeplitz.apply_normal(N, lam_d, km_d, cm2, v.reshape(bd.shape)).reshape(-1)
    ref_solve = ocg.cg(Ad, bd.reshape(-1), maxiter=20, reltol=0.0)
    b64 = toeplitz.backprojection(data, trj, cm, U=U)
    lam, km = toeplitz.toeplitz_kernel(N, trj, U)
    Af = lambda v: toeplitz.apply_normal(N, lam, km, cm2, v.reshape(b64.shape)).reshape(-1)
    ref = ocg.cg(Af, b64.reshape(-1), maxiter=20, reltol=0.0)
    print(f"small-case rel err: solve {rel_err(xr, ref_solve):.3e}, end-to-end {rel_err(xr, ref):.3e}",
          flush=True)
    for prec in ("mixed", "f32"):
        xp = msr.cg(A, b.reshape(-1), maxiter=20, reltol=0.0, precision=prec).cpu().numpy()
        print(f"small-case {prec}: solve {rel_err(xp, ref_solve):.3e}, end-to-end "
              f"{rel_err(xp, ref):.3e}", flush=True)
    assert rel_err(xr, ref_solve) < 1e-7
    assert rel_err(xr, ref) < RECON_TOL


def test_reference_radial_recon_and_layout_wrappers(torch, msr):
    """test/reconstruct_radial.jl and test/wrapper.jl restated: CG(20) close to
    the k-space-cropped truth; 3D and 4D array layouts give the same result."""
    N = 32
    K, C, nt, ncyc = 2, 4, 16, 6
    trj, U, cm, x = radial_case((N, N), K, C, nt, ncyc)
    data = toeplitz.forward(x.astype(np.complex128), trj, U, cm.astype(np.complex128)).astype(np.complex64)
    cml = [dev(torch, c) for c in cm]
    b = msr.calculate_backprojection(dev(torch, data), dev(torch, trj), cml, U=dev(torch, U))
    A = msr.NFFTNormalOp((N, N), dev(torch, trj), dev(torch, U), cmaps=cml)
    xr = msr.cg(A, b.reshape(-1), maxiter=20).cpu().numpy().reshape(x.shape)
    xc = np.fft.fftshift(np.fft.fft2(x.astype(np.complex128)), axes=(1, 2))
    iy, ix = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    xc[:, (ix - N / 2) ** 2 + (iy - N / 2) ** 2 > (N / 2) ** 2] = 0
    xc = np.fft.ifft2(np.fft.ifftshift(xc, axes=(1, 2)))
    assert rel_err(xr, xc) < 0.2
    # 4D layouts: trj (Nt, Ncyc, Nr, 2), data (C, Nt, Ncyc, Nr)
    trj4 = trj.reshape(nt, ncyc, 2 * N, 2)
    data4 = data.reshape(C, nt, ncyc, 2 * N)
    b4 = msr.calculate_backprojection(dev(torch, data4), dev(torch, trj4), cml, U=dev(torch, U))
    A4 = msr.NFFTNormalOp((N, N), dev(torch, trj4), dev(torch, U), cmaps=cml)
    xr4 = msr.cg(A4, b4.reshape(-1), maxiter=20).cpu().numpy().reshape(x.shape)
    assert rel_err(xr4, xr) < 1e-4


def test_backprojection_and_kernel_are_bit_reproducible(torch, msr):
    """2D spreading sums in a fixed order: repeated runs are bit-identical."""
    N = (48, 48)
    trj, U, cm, x = radial_case(N, 3, 3, 10, 3)
    rng = np.random.default_rng(5)
    data = (rng.normal(size=(3,) + trj.shape[:2]) + 1j * rng.normal(size=(3,) + trj.shape[:2])).astype(np.complex64)
    cml = [dev(torch, c) for c in cm]
    b1 = msr.calculate_backprojection(dev(torch, data), dev(torch, trj), cml, U=dev(torch, U)).cpu().numpy()
    b2 = msr.calculate_backprojection(dev(torch, data), dev(torch, trj), cml, U=dev(torch, U)).cpu().numpy()
    assert np.array_equal(b1, b2)
    l1 = msr.NFFTNormalOp(N, dev(torch, trj), dev(torch, U), cmaps=cml).prod.A.Λ.cpu().numpy()
    l2 = msr.NFFTNormalOp(N, dev(torch, trj), dev(torch, U), cmaps=cml).prod.A.Λ.cpu().numpy()
    assert np.array_equal(l1, l2)


def test_backprojection_gather_wide_offsets_same_bits(torch, msr, monkeypatch):
    """The matrix-core gather's 64-bit value offsets (taken above 2^25 samples)
    forced at a small size: the same bits as the 32-bit offsets."""
    N = (48, 40)
    trj, U, cm, x = radial_case(N, 3, 4, 10, 3)   # 12 channels: one group of 16
    rng = np.random.default_rng(6)
    data = (rng.normal(size=(4,) + trj.shape[:2]) + 1j * rng.normal(size=(4,) + trj.shape[:2])).astype(np.complex64)
    cml = [dev(torch, c) for c in cm]
    args = (dev(torch, data), dev(torch, trj), cml)
    b1 = msr.calculate_backprojection(*args, U=dev(torch, U)).cpu().numpy()
    monkeypatch.setenv("MSR_BP_WIDE_OFFSETS", "1")
    b2 = msr.calculate_backprojection(*args, U=dev(torch, U)).cpu().numpy()
    assert np.array_equal(b1, b2)
    assert np.linalg.norm(b1) > 0


@pytest.mark.parametrize("masked,C", [(False, 3), (True, 3), (False, 6), (True, 9), (False, 0)])
def test_forward_model_matches_oracle_and_is_adjoint(torch, msr, masked, C):
    """Forward (type-2) model vs the FP64 oracle, and <A x, y> = <x, A^H y> with
    the device backprojection as A^H.  Coil counts cover whole and partial
    groups of the interpolation; C = 0 is no coil maps (one unit coil)."""
    N = (32, 24)
    trj, U, cm, x = radial_case(N, 3, max(C, 1), 8, 3)
    if C == 0:
        cm = np.ones((1,) + N[::-1], dtype=np.complex64)
    mask = None
    if masked:
        mask = np.ones(trj.shape[:2], dtype=bool)
        mask[2, 10:50] = False
    md = None if mask is None else dev(torch, mask)
    cml = [dev(torch, c) for c in cm]
    y = msr.forward(dev(torch, x), dev(torch, trj), cml if C else None, U=dev(torch, U),
                    sample_mask=md).cpu().numpy()
    ref = toeplitz.forward(x.astype(np.complex128), trj, U, cm.astype(np.complex128), sample_mask=mask)
    print(f"forward rel err {rel_err(y, ref):.3e}", flush=True)
    assert rel_err(y, ref) < 1e-5
    rng = np.random.default_rng(9)
    d = (rng.normal(size=y.shape) + 1j * rng.normal(size=y.shape)).astype(np.complex64)
    if mask is not None:
        d[:, ~mask] = 0
    b = msr.calculate_backprojection(dev(torch, d), dev(torch, trj), cml, U=dev(torch, U),
                                     sample_mask=md).cpu().numpy()
    lhs = np.vdot(y.astype(np.complex128), d.astype(np.complex128))
    rhs = np.vdot(x.astype(np.complex128), b.astype(np.complex128))
    assert abs(lhs - rhs) < 1e-5 * abs(lhs)


@pytest.mark.parametrize("three_d", [False, True])
def test_ragged_frames_empty_and_single_sample(torch, msr, three_d):
    """Sample masks that empty whole frames and leave single samples in others
    (ragged per-frame counts, nsamp_t = 0 and 1): kmask, kernel,
    backprojection and forward model against the oracle, and a CG solve."""
    from oracle import cg as ocg
    if three_d:
        N = (16, 16, 16)
        synth = msr.synth
        nt, nspk = 6, 4
        trj = synth.traj_kooshball_goldenratio(2 * max(N), nspk, nt).astype(np.float32)
        U = synth.t2_basis(nt, 2, seed=1)
        cm = synth.gaussian_coil_maps(N, 2, seed=1)
    else:
        N = (24, 32)
        nt, nspk = 7, 3
        trj, U, cm, _ = radial_case(N, 3, 2, nt, nspk, seed=2)
    mask = np.ones(trj.shape[:2], dtype=bool)
    mask[0] = False                    # frame 0 empty
    mask[3] = False
    mask[3, 17] = True                 # frame 3: one sample
    mask[nt - 1, ::2] = False          # last frame: every other sample
    rng = np.random.default_rng(3)
    C = len(cm)
    data = (rng.normal(size=(C,) + trj.shape[:2]) + 1j * rng.normal(size=(C,) + trj.shape[:2])).astype(np.complex64)
    cml = [dev(torch, c) for c in cm]
    A = msr.NFFTNormalOp(N, dev(torch, trj), dev(torch, U), cmaps=cml, sample_mask=dev(torch, mask))
    lam_ref, km_ref = toeplitz.toeplitz_kernel(N, trj, U, sample_mask=mask)
    assert np.array_equal(A.prod.A.kmask_indcs.cpu().numpy(), km_ref + 1)
    e_k = rel_err(A.prod.A.Λ.cpu().numpy(), lam_ref)
    b = msr.calculate_backprojection(dev(torch, data), dev(torch, trj), cml, U=dev(torch, U),
                                     sample_mask=dev(torch, mask))
    b_ref = toeplitz.backprojection(data, trj, cm, U=U, sample_mask=mask)
    e_b = rel_err(b.cpu().numpy(), b_ref)
    x = (rng.normal(size=b_ref.shape) + 1j * rng.normal(size=b_ref.shape)).astype(np.complex64)
    y = msr.forward(dev(torch, x), dev(torch, trj), cmaps=cml, U=dev(torch, U),
                    sample_mask=dev(torch, mask)).cpu().numpy()
    y_ref = toeplitz.forward(x, trj, U, cm, sample_mask=mask)
    e_f = rel_err(y, y_ref)
    assert np.all(y[:, 0] == 0)       # the empty frame stays zero
    xs = msr.cg(A, b.reshape(-1), maxiter=8, reltol=0.0).cpu().numpy()
    op = toeplitz.NormalOp(N, lam_ref, km_ref, cm)
    xr = ocg.cg(op, b_ref.reshape(-1).astype(np.complex128), 8, reltol=0.0)
    e_s = rel_err(xs, xr)
    print(f"ragged frames 3d={three_d}: kernel {e_k:.2e} bp {e_b:.2e} forward {e_f:.2e} "
          f"solve {e_s:.2e}", flush=True)
    assert e_k < KERNEL_TOL and e_b < 2e-6 and e_f < 2e-6 and e_s < RECON_TOL
