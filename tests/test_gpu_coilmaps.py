"""Coil-map calibration front end on the device against the FP64 oracle
(oracle/coilmaps.py): calibration region bit-exact, calibration k-space to
FP32 rounding, coil-wise reconstruction (all coils in lockstep on one shared
kernel) within 1e-4 of the FP64 restatement, and recovery of coil images
from exact calibration data.  The ESPIRiT step is the caller's; the plumbing
around it (layouts in and out) is checked with a stand-in."""
import numpy as np
import pytest

from conftest import load_package, rel_err
from oracle import coilmaps as ocm

pytestmark = pytest.mark.gpu
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


def case2d(N=(64, 64), nt=12, nspk=3, C=3, seed=0):
    synth = load_package().synth
    trj = synth.traj_2d_radial_goldenratio(2 * max(N), nspk, nt)
    rng = np.random.default_rng(seed)
    data = (rng.normal(size=(C,) + trj.shape[:2]) + 1j * rng.normal(size=(C,) + trj.shape[:2]))
    return trj, data.astype(np.complex64)


def case3d(N=(32, 32, 32), nt=6, nspk=8, C=2, seed=0):
    synth = load_package().synth
    trj = synth.traj_kooshball_goldenratio(2 * max(N), nspk, nt).astype(np.float32)
    rng = np.random.default_rng(seed)
    data = (rng.normal(size=(C,) + trj.shape[:2]) + 1j * rng.normal(size=(C,) + trj.shape[:2]))
    return trj, data.astype(np.complex64)


@pytest.mark.parametrize("three_d,masked", [(False, False), (False, True), (True, True)])
def test_calibration_region_is_bit_exact(torch, msr, three_d, masked):
    N, cs = ((32, 32, 32), (16, 16, 8)) if three_d else ((64, 48), (16, 24))
    trj, _ = case3d(N) if three_d else case2d(N)
    mask = None
    if masked:
        mask = np.random.default_rng(1).uniform(size=trj.shape[:2]) < 0.7
    tc, mc = msr.calibration_region(dev(torch, trj), N, cs,
                                    None if mask is None else dev(torch, mask))
    rt, rm = ocm.calib_select(trj, N, cs, mask)
    assert np.array_equal(tc.cpu().numpy(), rt)
    assert np.array_equal(mc.cpu().numpy(), rm)
    assert 0 < rm.sum() < rm.size
    with pytest.raises(AssertionError):
        msr.calibration_region(dev(torch, trj), N, (7,) * len(N))  # not 2,3,5-smooth


@pytest.mark.parametrize("shape", [(32, 32), (16, 24), (16, 24, 32)])
def test_calibration_kspace_matches_centred_fft(torch, msr, shape):
    rng = np.random.default_rng(2)
    C = 3
    x = (rng.normal(size=(C,) + shape[::-1]) + 1j * rng.normal(size=(C,) + shape[::-1])).astype(np.complex64)
    k = msr.calibration_kspace(dev(torch, x)).cpu().numpy()
    ref = ocm.calib_kspace(x, len(shape))
    print(f"calib k-space {shape}: rel err {rel_err(k, ref):.2e}", flush=True)
    assert rel_err(k, ref) < 2e-6


def test_calibration_kspace_rejects_lengths_without_fft(torch, msr):
    x = torch.zeros((2, 31, 31), dtype=torch.complex64, device="cuda")  # prime: no FFT
    with pytest.raises(NotImplementedError):
        msr.calibration_kspace(x)


# A complex basis column gives a complex K = 1 kernel whose imaginary part (from
# the unpaired r = -N row of the PSF, through the kmask truncation) makes the
# operator slightly non-Hermitian; CG on it then amplifies the ComplexF32
# rounding of the stored kernel (the FP64 oracle itself moves 5% at 15
# iterations when only its kernel is rounded to ComplexF32), so that case is
# compared over the iterations where CG is still stable.  The calibration
# default, U = ones, gives a real kernel and a Hermitian operator.
@pytest.mark.parametrize("basis,niter", [("ones", 20), ("ones", 60), ("complex", 8)])
def test_coilwise_matches_fp64_oracle(torch, msr, basis, niter):
    N, cs = (64, 64), (16, 16)
    trj, data = case2d(N)
    tc, mc = ocm.calib_select(trj, N, cs)
    U = None
    if basis == "complex":
        rng = np.random.default_rng(3)
        U = (rng.normal(size=(2, trj.shape[0])) + 1j * rng.normal(size=(2, trj.shape[0]))).astype(np.complex64)
    x = msr.reconstruct_coilwise(dev(torch, data), dev(torch, tc), cs,
                                 U=None if U is None else dev(torch, U),
                                 sample_mask=dev(torch, mc), Niter_cg=niter).cpu().numpy()
    ref = ocm.reconstruct_coilwise(data, tc, cs, U=U, sample_mask=mc, niter=niter)
    err = rel_err(x, ref)
    per_coil = [rel_err(x[c], ref[c]) for c in range(x.shape[0])]
    print(f"coilwise {basis} niter={niter}: rel err {err:.2e}, per coil "
          f"{', '.join(f'{e:.1e}' for e in per_coil)}", flush=True)
    assert err <= RECON_TOL and max(per_coil) <= RECON_TOL
    # the reference's ComplexF32 arithmetic stays close as well
    x32 = msr.reconstruct_coilwise(dev(torch, data), dev(torch, tc), cs,
                                   U=None if U is None else dev(torch, U),
                                   sample_mask=dev(torch, mc), Niter_cg=niter,
                                   precision="f32").cpu().numpy()
    print(f"  f32 arithmetic: rel err {rel_err(x32, ref):.2e}", flush=True)
    assert rel_err(x32, ref) < 1e-2


def test_coilwise_3d_matches_fp64_oracle(torch, msr):
    N, cs = (32, 32, 32), (16, 16, 16)
    trj, data = case3d(N)
    tc, mc = ocm.calib_select(trj, N, cs)
    x = msr.reconstruct_coilwise(dev(torch, data), dev(torch, tc), cs, sample_mask=dev(torch, mc),
                                 Niter_cg=10).cpu().numpy()
    ref = ocm.reconstruct_coilwise(data, tc, cs, sample_mask=mc, niter=10)
    print(f"coilwise 3d: rel err {rel_err(x, ref):.2e}", flush=True)
    assert rel_err(x, ref) <= RECON_TOL


def test_coilwise_recovers_coil_images(torch, msr):
    N, cs, C = (64, 64), (16, 16), 4
    trj, _ = case2d(N, nt=16, C=C)
    tc, mc = ocm.calib_select(trj, N, cs)
    rng = np.random.default_rng(4)
    xt = rng.normal(size=(C, 16, 16)) + 1j * rng.normal(size=(C, 16, 16))
    ry, rx = np.meshgrid(np.arange(16) - 8, np.arange(16) - 8, indexing="ij")
    r = np.stack([rx.reshape(-1), ry.reshape(-1)], 1)
    E = np.exp(-2j * np.pi * tc[mc].astype(np.float64) @ r.T)
    data = np.zeros((C,) + trj.shape[:2], dtype=np.complex64)
    for c in range(C):
        data[c][mc] = (E @ xt[c].reshape(-1)).astype(np.complex64)
    x = msr.reconstruct_coilwise(dev(torch, data), dev(torch, tc), cs, sample_mask=dev(torch, mc),
                                 Niter_cg=100).cpu().numpy()
    print(f"coilwise recovery: rel err {rel_err(x, xt):.2e}", flush=True)
    assert rel_err(x, xt) < 1e-4


def test_calculate_coil_maps_plumbing(torch, msr):
    N = (64, 64)
    trj, data = case2d(N, C=2)
    cs = msr.default_calib_size(N)
    assert cs == (32, 32)
    with pytest.raises(NotImplementedError):
        msr.calculate_coil_maps(dev(torch, data), dev(torch, trj), N)
    seen = {}

    def espirit(kbp, img_shape, kernel_size, eigThresh_1, eigThresh_2, nmaps):
        # stand-in: records its input, returns (img..., C, nmaps) maps whose
        # values name their position
        seen["kbp"] = kbp
        seen["args"] = (img_shape, kernel_size, eigThresh_1, eigThresh_2, nmaps)
        C = kbp.shape[-1]
        ix, iy = np.meshgrid(np.arange(img_shape[0]), np.arange(img_shape[1]), indexing="ij")
        return np.stack([np.stack([(ix + 1j * iy) * (c + 1) * (n + 1) for c in range(C)], -1)
                         for n in range(nmaps)], -1)

    maps = msr.calculate_coil_maps(dev(torch, data), dev(torch, trj), N, espirit=espirit,
                                   nmaps=2, Niter_cg=10)
    tc, mc = ocm.calib_select(trj, N, cs)
    xr = ocm.reconstruct_coilwise(data, tc, cs, sample_mask=mc, niter=10)
    kref = ocm.calib_kspace(xr, 2)  # (C, 32, 32) row-major
    kbp = seen["kbp"]
    assert kbp.shape == (32, 32, 2)  # Julia (calib..., C)
    assert rel_err(np.moveaxis(kbp, -1, 0).transpose(0, 2, 1), kref) <= RECON_TOL
    assert seen["args"] == (N, (6, 6), 0.01, 0.9, 2)
    assert len(maps) == 4  # coil fastest, then map
    for i, m in enumerate(maps):
        c, n = i % 2, i // 2
        assert tuple(m.shape) == N[::-1]
        mm = m.cpu().numpy()
        assert mm[5, 3] == np.complex64((3 + 5j) * (c + 1) * (n + 1))  # [y, x] = Julia [x, y]
