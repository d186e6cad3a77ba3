"""GPU parity of the 3D path (five-pass fused apply, 3D kernel build, 3D
backprojection, CG) against the FP64 oracle on koosh-ball trajectories
(src/Trajectories.jl:51-60).  Tolerances as in the 2D tests: apply 2e-6,
kernel / backprojection 1e-5, solve on the device-built operator 1e-4."""
import numpy as np
import pytest

from conftest import load_package, rel_err
from oracle import cg as ocg
from oracle import toeplitz

pytestmark = pytest.mark.gpu


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


@pytest.mark.parametrize("N,K,C,real,identity", [
    ((16, 16, 16), 3, 3, False, False),
    ((16, 8, 24), 2, 2, True, False),
    ((8, 16, 16), 4, 1, False, True),
    ((32, 32, 32), 4, 4, False, False),
])
def test_apply3d_matches_oracle(torch, msr, N, K, C, real, identity):
    print(f"3d apply N={N} K={K} C={C} real={real}", flush=True)
    rng = np.random.default_rng(abs(hash((N, K, C))) % 2 ** 32)
    G = int(np.prod([2 * v for v in N]))
    km = np.sort(rng.choice(G, size=int(0.7 * G), replace=False)).astype(np.int64)
    P = K * (K + 1) // 2
    lam = rng.normal(size=(km.size, P))
    lam = (lam if real else lam + 1j * rng.normal(size=(km.size, P))).astype(np.float32 if real else np.complex64)
    cm = None if identity else msr.synth.gaussian_coil_maps(N, C, seed=3)
    x = (rng.normal(size=(K,) + N[::-1]) + 1j * rng.normal(size=(K,) + N[::-1])).astype(np.complex64)
    A = msr.NFFTNormalOp(N, dev(torch, lam), dev(torch, km + 1),
                         cmaps=None if cm is None else [dev(torch, c) for c in cm])
    xd = dev(torch, x).reshape(-1)
    out = (A @ xd).cpu().numpy().reshape(x.shape)
    ref = toeplitz.apply_normal(N, lam.astype(np.float64 if real else np.complex128), km,
                                None if cm is None else cm.astype(np.complex128), x.astype(np.complex128))
    print(f"3d apply rel err {rel_err(out, ref):.3e}", flush=True)
    assert rel_err(out, ref) < 2e-6
    y = torch.empty_like(xd)
    A.A.mul_(y, xd, 0.5, 0.0, percoil=True)
    assert rel_err(y.cpu().numpy().reshape(x.shape), 0.5 * ref) < 2e-6


def kooshball_case(N, K, C, nt, nspk, seed=0):
    synth = load_package().synth
    trj = synth.traj_kooshball_goldenratio(2 * max(N), nspk, nt).astype(np.float32)
    U = synth.t2_basis(nt, K, seed=seed)
    cm = synth.gaussian_coil_maps(N, C, seed=seed)
    x = synth.ellipse_phantom(N, K, seed=seed)
    return trj, U, cm, x


def test_kernel3d_and_backprojection3d(torch, msr):
    N = (16, 16, 16)
    trj, U, cm, x = kooshball_case(N, 2, 3, 8, 6)
    cml = [dev(torch, c) for c in cm]
    A = msr.NFFTNormalOp(N, dev(torch, trj), dev(torch, U), cmaps=cml)
    lam_ref, km_ref = toeplitz.toeplitz_kernel(N, trj, U)
    assert np.array_equal(A.prod.A.kmask_indcs.cpu().numpy(), km_ref + 1)
    lam = A.prod.A.Λ.cpu().numpy()
    print(f"3d kernel rel err {rel_err(lam, lam_ref):.3e}", flush=True)
    assert rel_err(lam, lam_ref) < 3e-7  # FP64 build, W = 8 truncation
    rng = np.random.default_rng(4)
    data = (rng.normal(size=(3,) + trj.shape[:2]) + 1j * rng.normal(size=(3,) + trj.shape[:2])).astype(np.complex64)
    for dcf in ("none", "radial_3D"):
        b = msr.calculate_backprojection(dev(torch, data), dev(torch, trj), cml, U=dev(torch, U),
                                         density_compensation=dcf).cpu().numpy()
        ref = toeplitz.backprojection(data, trj, cm, U=U, density_compensation=dcf)
        print(f"3d bp ({dcf}) rel err {rel_err(b, ref):.3e}", flush=True)
        assert rel_err(b, ref) < 1e-5
    pc = msr.calculate_backprojection(dev(torch, data), dev(torch, trj), N, U=dev(torch, U)).cpu().numpy()
    assert rel_err(pc, toeplitz.backprojection(data, trj, N, U=U)) < 1e-5


@pytest.mark.parametrize("N,K,C,dcf,masked", [
    ((9, 12, 15), 2, 3, "radial_3D", False),  # odd N: 2N = 18, 30 leave partial 4-cell blocks
    ((20, 16, 24), 3, 2, "radial_2D", True),  # masked samples, the radial_2D weight in 3D
    ((16, 16, 16), 4, 5, "none", False),      # 20 channels: groups of 16 and 4 (padded)
    ((24, 24, 24), 1, 1, "none", False),      # one channel: the single-channel gather
])
def test_backprojection3d_gather_matches_oracle(torch, msr, N, K, C, dcf, masked):
    """The 3D output-block gather (4 x 4 x 4 blocks, split dense centre blocks,
    grids not a multiple of the block) against the FP64 NUDFT-checked oracle,
    coil-combined and per coil, and bit-identical from call to call."""
    print(f"3d bp gather N={N} K={K} C={C} dcf={dcf} masked={masked}", flush=True)
    trj, U, cm, x = kooshball_case(N, K, C, 12, 8, seed=N[0])
    rng = np.random.default_rng(N[2])
    data = (rng.normal(size=(C,) + trj.shape[:2]) + 1j * rng.normal(size=(C,) + trj.shape[:2])).astype(np.complex64)
    mask = None
    if masked:
        mask = rng.uniform(size=trj.shape[:2]) < 0.7
    args = dict(U=dev(torch, U), density_compensation=dcf,
                sample_mask=None if mask is None else dev(torch, mask))
    cml = [dev(torch, c) for c in cm]
    traj = msr.plan_trajectory(dev(torch, trj), None if mask is None else dev(torch, mask))
    b1 = msr.calculate_backprojection(dev(torch, data), traj, cml, U=dev(torch, U),
                                      density_compensation=dcf)
    b2 = msr.calculate_backprojection(dev(torch, data), traj, cml, U=dev(torch, U),
                                      density_compensation=dcf)
    assert torch.equal(b1, b2)
    ref = toeplitz.backprojection(data, trj, cm, U=U, sample_mask=mask, density_compensation=dcf)
    e = rel_err(b1.cpu().numpy(), ref)
    print(f"  combined rel err {e:.3e}", flush=True)
    assert e < 1e-5
    pc = msr.calculate_backprojection(dev(torch, data), dev(torch, trj), N, **args).cpu().numpy()
    ref_pc = toeplitz.backprojection(data, trj, N, U=U, sample_mask=mask, density_compensation=dcf)
    assert rel_err(pc, ref_pc) < 1e-5


def test_recon3d_solve_matches_fp64(torch, msr):
    """The kmask-truncated Toeplitz operator is indefinite (its smallest
    eigenvalue is about -2e-3 of the largest on this 16^3 koosh-ball case,
    computed with ARPACK on the FP64 oracle), so CG amplifies the rounding of
    each apply more the longer it runs: an FP32 apply reaches ~2e-4 after 15
    iterations.  The default FP64 apply holds the solve to 1e-7 at 8 and 15
    iterations; the FP32-apply ("mixed") result is printed."""
    N = (16, 16, 16)
    trj, U, cm, x = kooshball_case(N, 2, 4, 10, 8, seed=2)
    data = toeplitz.forward(x.astype(np.complex128), trj, U, cm.astype(np.complex128)).astype(np.complex64)
    cml = [dev(torch, c) for c in cm]
    b = msr.calculate_backprojection(dev(torch, data), dev(torch, trj), cml, U=dev(torch, U))
    A = msr.NFFTNormalOp(N, dev(torch, trj), dev(torch, U), cmaps=cml)
    lam_d = A.prod.A.Λ.cpu().numpy().astype(np.complex128)
    km_d = A.prod.A.kmask_indcs.cpu().numpy() - 1
    op = toeplitz.NormalOp(N, lam_d, km_d, cm.astype(np.complex128))
    bd = b.cpu().numpy().astype(np.complex128).reshape(-1)
    for mode in ("device", "graph"):
        for iters in (8, 15):
            xr = msr.cg(A, b.reshape(-1), maxiter=iters, reltol=0.0, mode=mode).cpu().numpy()
            ref = ocg.cg(op, bd, maxiter=iters, reltol=0.0)
            xm = msr.cg(A, b.reshape(-1), maxiter=iters, reltol=0.0, mode=mode,
                        precision="mixed").cpu().numpy()
            print(f"3d solve ({mode}, {iters} it) rel err {rel_err(xr, ref):.3e} "
                  f"(FP32 apply {rel_err(xm, ref):.3e})", flush=True)
            assert rel_err(xr, ref) < 1e-7


def test_3d_spreading_atomic_and_slab_paths_agree(torch, msr, monkeypatch):
    """3D tile spreading stores padded tiles to a slab summed by a gather, or,
    when one channel group's slab exceeds the budget (forced here with
    MSR_SPREAD3D_SLAB_BUDGET=1), adds them into the grid with atomics: kernel
    and backprojection agree between the two paths and with the oracle."""
    N = (16, 16, 16)
    trj, U, cm, x = kooshball_case(N, 2, 3, 6, 5)
    cml = [dev(torch, c) for c in cm]
    rng = np.random.default_rng(8)
    data = (rng.normal(size=(3,) + trj.shape[:2]) + 1j * rng.normal(size=(3,) + trj.shape[:2])).astype(np.complex64)

    def run():
        A = msr.NFFTNormalOp(N, dev(torch, trj), dev(torch, U), cmaps=cml)
        b = msr.calculate_backprojection(dev(torch, data), dev(torch, trj), cml, U=dev(torch, U))
        return A.prod.A.Λ.cpu().numpy(), b.cpu().numpy()

    lam_s, b_s = run()
    monkeypatch.setenv("MSR_SPREAD3D_SLAB_BUDGET", "1")
    lam_a, b_a = run()
    monkeypatch.delenv("MSR_SPREAD3D_SLAB_BUDGET")
    lam_ref, _ = toeplitz.toeplitz_kernel(N, trj, U)
    b_ref = toeplitz.backprojection(data, trj, cm, U=U)
    print(f"3d paths: kernel slab/atomic {rel_err(lam_s, lam_a):.1e}, bp {rel_err(b_s, b_a):.1e}; "
          f"vs oracle {rel_err(lam_a, lam_ref):.1e} {rel_err(b_a, b_ref):.1e}", flush=True)
    assert rel_err(lam_s, lam_a) < 1e-7 and rel_err(b_s, b_a) < 1e-6
    assert rel_err(lam_a, lam_ref) < 3e-7 and rel_err(b_a, b_ref) < 2e-6


def test_backprojection3d_gather_wide_offsets_same_bits(torch, msr, monkeypatch):
    """The 3D matrix-core gather with its 64-bit value offsets (taken above 2^25
    samples) forced at a small size: the same bits as the 32-bit offsets."""
    N = (16, 16, 12)
    trj, U, cm, x = kooshball_case(N, 3, 4, 6, 5)   # 12 channels: one group of 16
    cml = [dev(torch, c) for c in cm]
    rng = np.random.default_rng(9)
    data = (rng.normal(size=(4,) + trj.shape[:2]) + 1j * rng.normal(size=(4,) + trj.shape[:2])).astype(np.complex64)
    args = (dev(torch, data), dev(torch, trj), cml)
    b1 = msr.calculate_backprojection(*args, U=dev(torch, U)).cpu().numpy()
    monkeypatch.setenv("MSR_BP_WIDE_OFFSETS", "1")
    b2 = msr.calculate_backprojection(*args, U=dev(torch, U)).cpu().numpy()
    assert np.array_equal(b1, b2)
    assert np.linalg.norm(b1) > 0


def test_recon3d_compact_solve(torch, msr):
    """The f64c precision in 3D: FP64 arithmetic with the apply's T1 / T2
    intermediates stored in 8 bytes (four storage roundings of 2^-27 per
    apply) against the FP64 oracle's CG on the same operator, 15 iterations."""
    N = (16, 16, 16)
    trj, U, cm, x = kooshball_case(N, 2, 4, 10, 8, seed=5)
    cml = [dev(torch, c) for c in cm]
    A = msr.NFFTNormalOp(N, dev(torch, trj), dev(torch, U), cmaps=cml)
    rng = np.random.default_rng(6)
    data = (rng.normal(size=(4,) + trj.shape[:2]) + 1j * rng.normal(size=(4,) + trj.shape[:2])).astype(np.complex64)
    b = msr.calculate_backprojection(dev(torch, data), dev(torch, trj), cml, U=dev(torch, U))
    bd = b.reshape(-1)
    lam = A.prod.A.Λ.cpu().numpy().astype(np.complex128)
    km = A.prod.A.kmask_indcs.cpu().numpy() - 1
    bv = b.cpu().numpy()
    Af = lambda v: toeplitz.apply_normal(N, lam, km, cm.astype(np.complex128),
                                         v.reshape(bv.shape)).reshape(-1)
    ref = ocg.cg(Af, bv.reshape(-1).astype(np.complex128), maxiter=15, reltol=0.0)
    for mode in ("device", "graph"):
        xg = msr.cg(A, bd, maxiter=15, reltol=0.0, mode=mode, precision="f64c")
        e = rel_err(xg.cpu().numpy(), ref)
        print(f"3d solve f64c ({mode}) rel err {e:.3e}", flush=True)
        assert e < 2e-5



@pytest.mark.parametrize("N,K,C", [((16, 16, 16), 3, 3), ((24, 20, 12), 2, 2)])
def test_compact_column_pairs_same_bits(torch, msr, monkeypatch, N, K, C):
    """The 3D apply's y-column passes over 8-byte values (compact ComplexF64,
    ComplexF32) load and store two adjacent columns per 16-byte access, and its
    x-row passes A and E two consecutive compact values;
    MSR_COLS_SINGLE=1 takes one value per access.  The decode / encode of each
    value is the same, so the compact apply, a compact device CG and the
    ComplexF32 apply give the same bits either way."""
    from test_gpu_apply import make_case
    print(f"compact column pairs N={N} K={K} C={C}", flush=True)
    rng = np.random.default_rng(21)
    lam, km, cm, x = make_case(rng, N, K, C)
    A = msr.NFFTNormalOp(N, dev(torch, lam), dev(torch, km + 1), cmaps=[dev(torch, c) for c in cm])
    xd64 = dev(torch, x.astype(np.complex128) + 1e-3j * rng.normal(size=x.shape)).reshape(-1)
    xd = dev(torch, x).reshape(-1)

    def run():
        y = torch.empty_like(xd64)
        A.A.mul_f64_(y, xd64, compact=True)
        s = msr.cg(A, xd, maxiter=5, reltol=0.0, mode="device", precision="f64c")
        return y.cpu().numpy(), s.cpu().numpy(), (A @ xd).cpu().numpy()

    paired = run()
    monkeypatch.setenv("MSR_COLS_SINGLE", "1")
    single = run()
    for name, a, b in zip(("apply", "cg", "f32 apply"), paired, single):
        print(f"  {name}: max |diff| {np.max(np.abs(a - b)):.3e}", flush=True)
        assert np.array_equal(a, b), name
    assert np.linalg.norm(paired[0]) > 0


@pytest.mark.parametrize("dim", [2, 3])
def test_backprojection_column_pairs_same_bits(torch, msr, monkeypatch, dim):
    """The backprojection's ComplexF32 line passes with two adjacent columns
    (column passes) or two consecutive elements (row passes) per 16-byte
    access and with one value per access (MSR_COLS_SINGLE=1): the same bits,
    in 2D and 3D."""
    if dim == 3:
        N = (16, 16, 12)
        trj, U, cm, _ = kooshball_case(N, 3, 4, 6, 5)
    else:
        N = (40, 36)
        trj = msr.synth.traj_2d_radial_goldenratio(80, 3, 10).astype(np.float32)
        U = msr.synth.t2_basis(10, 3)
        cm = msr.synth.gaussian_coil_maps(N, 4)
    cml = [dev(torch, c) for c in cm]
    rng = np.random.default_rng(13)
    data = (rng.normal(size=(4,) + trj.shape[:2]) + 1j * rng.normal(size=(4,) + trj.shape[:2])).astype(np.complex64)
    args = (dev(torch, data), dev(torch, trj), cml)
    b1 = msr.calculate_backprojection(*args, U=dev(torch, U)).cpu().numpy()
    monkeypatch.setenv("MSR_COLS_SINGLE", "1")
    b2 = msr.calculate_backprojection(*args, U=dev(torch, U)).cpu().numpy()
    print(f"bp column pairs {dim}D: max |diff| {np.max(np.abs(b1 - b2)):.3e}", flush=True)
    assert np.array_equal(b1, b2)
    assert np.linalg.norm(b1) > 0
