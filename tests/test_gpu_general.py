"""Grid and coefficient-count generality: the line-kernel Toeplitz apply that
serves every shape outside the fused passes (2N any 2,3,5-smooth length up to
2048, 2Ny above 512, K above 8, 3D with a non-fused x or z length), and the
line FFTs of the backprojection, kernel construction and Cartesian operator at
those lengths, against the FP64 oracle.

Tolerances (relative L2), as for the fused passes: one ComplexF32 apply 2e-6,
one ComplexF64 apply 1e-12 (the oracle on the same FP32 kernel and maps),
kernel 3e-7 and backprojection 1e-5 (FP32 NUFFT), kmask bit-exact, CG in FP64
1e-7 against the oracle's CG on the device-built kernel."""
import os

import numpy as np
import pytest

from conftest import load_package, rel_err
from oracle import cartesian as ocart
from oracle import cg as ocg
from oracle import toeplitz
from test_gpu_apply import make_case
from test_oracle_cartesian import traj_cartesian

pytestmark = pytest.mark.gpu

APPLY_TOL = 2e-6
F64_TOL = 1e-12
KERNEL_TOL = 3e-7
NUFFT_TOL = 1e-5


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


def make_op(torch, msr, N, lam, km, cm):
    cml = None if cm is None else [dev(torch, c) for c in cm]
    return msr.NFFTNormalOp(N, dev(torch, lam), dev(torch, km + 1), cmaps=cml)


def oracle_apply(N, lam, km, cm, x):
    lam = lam.astype(np.float64 if lam.dtype == np.float32 else np.complex128)
    return toeplitz.apply_normal(N, lam, km, None if cm is None else cm.astype(np.complex128),
                                 x.astype(np.complex128))


# (N, K, C, real kernel, identity maps): shapes and K that no fused pass takes
LINE_CASES = [
    ((72, 72), 3, 4, False, False),     # 2N = 144
    ((100, 60), 4, 3, True, False),     # 200 x 120
    ((120, 40), 2, 2, False, False),    # 240 x 80
    ((240, 24), 3, 2, False, False),    # 480 x 48
    ((36, 20), 2, 3, False, True),      # 72 x 40, identity maps
    ((20, 300), 2, 2, False, False),    # 2Ny = 600 above 512
    ((45, 27), 2, 2, False, False),     # 90 x 54: odd factors, non-exact plans
    ((32, 32), 10, 2, False, False),    # K = 10
    ((16, 16), 12, 1, True, False),     # K = 12, real kernel
    ((12, 10, 9), 3, 2, False, False),  # 3D, 2Nz = 18
    ((36, 16, 8), 2, 2, True, False),   # 3D, 2Nx = 72
    ((16, 16, 16), 9, 2, False, False), # 3D, K = 9
]


@pytest.mark.parametrize("N,K,C,real,identity", LINE_CASES)
def test_line_apply_matches_oracle(torch, msr, N, K, C, real, identity):
    print(f"line apply N={N} K={K} C={C} real={real} identity={identity}", flush=True)
    rng = np.random.default_rng(abs(hash((N, K, C, real, identity))) % 2 ** 32)
    lam, km, cm, x = make_case(rng, N, K, C, real, identity)
    A = make_op(torch, msr, N, lam, km, cm)
    xd = dev(torch, x).reshape(-1)
    ref = oracle_apply(N, lam, km, cm, x)
    out = (A @ xd).reshape(x.shape).cpu().numpy()
    e = rel_err(out, ref)
    print(f"  f32 apply rel err {e:.2e}", flush=True)
    assert e < APPLY_TOL
    # the per-coil order
    y = torch.empty_like(xd)
    A.A.mul_(y, xd, 1.0, 0.0, percoil=True)
    assert rel_err(y.reshape(x.shape).cpu().numpy(), ref) < APPLY_TOL
    # ComplexF64 (and the compact mode, which this path runs in ComplexF64)
    x64 = x.astype(np.complex128) + 1e-3j * rng.normal(size=x.shape)
    xd64 = dev(torch, x64).reshape(-1)
    ref64 = oracle_apply(N, lam, km, cm, x64)
    for compact in (False, True):
        y64 = torch.empty_like(xd64)
        A.A.mul_f64_(y64, xd64, compact=compact)
        e64 = rel_err(y64.reshape(x.shape).cpu().numpy(), ref64)
        print(f"  f64 apply (compact={compact}) rel err {e64:.2e}", flush=True)
        assert e64 < F64_TOL


def test_line_apply_matches_fused_apply(torch, msr):
    """MSR_APPLY_LINES=1 at operator creation sends a shape the fused passes
    take to the line-kernel apply: both paths agree (FP64 to 1e-12, FP32 to
    the FP32 apply tolerance) and differ in rounding, which shows the switch
    took effect."""
    rng = np.random.default_rng(5)
    N, K, C = (64, 48), 5, 4
    lam, km, cm, x = make_case(rng, N, K, C)
    A = make_op(torch, msr, N, lam, km, cm)
    os.environ["MSR_APPLY_LINES"] = "1"
    try:
        L = make_op(torch, msr, N, lam, km, cm)
    finally:
        del os.environ["MSR_APPLY_LINES"]
    xd = dev(torch, x).reshape(-1)
    ya, yl = (A @ xd).cpu().numpy(), (L @ xd).cpu().numpy()
    print(f"fused vs lines f32: {rel_err(yl, ya):.2e}", flush=True)
    assert rel_err(yl, ya) < APPLY_TOL
    assert not np.array_equal(yl, ya)
    x64 = dev(torch, x.astype(np.complex128)).reshape(-1)
    a64, l64 = torch.empty_like(x64), torch.empty_like(x64)
    A.A.mul_f64_(a64, x64)
    L.A.mul_f64_(l64, x64)
    assert rel_err(l64.cpu().numpy(), a64.cpu().numpy()) < F64_TOL
    # device CG (graph capture of the line passes) in FP64 on both paths
    xa = msr.cg(A, xd, maxiter=8, reltol=0.0, mode="graph", precision="f64")
    xl = msr.cg(L, xd, maxiter=8, reltol=0.0, mode="graph", precision="f64")
    e = rel_err(xl.cpu().numpy(), xa.cpu().numpy())
    print(f"fused vs lines CG f64: {e:.2e}", flush=True)
    assert e < 1e-7


def radial(N, K, C, nt, nspk, seed=0):
    synth = load_package().synth
    trj = synth.traj_2d_radial_goldenratio(2 * max(N), nspk, nt)
    U = synth.t2_basis(nt, K, seed=seed)
    cm = synth.gaussian_coil_maps(N, C, seed=seed)
    return trj, U, cm


@pytest.mark.parametrize("N", [(72, 72), (100, 100), (120, 120), (240, 240)])
def test_reconstruction_at_general_sizes(torch, msr, N):
    """Kernel construction (4N FFTs: 288 ... 960), backprojection (2N FFTs) and
    an FP64 device CG on the line-kernel apply, each against the oracle."""
    K, C, nt, nspk = 3, 3, 8, 3
    print(f"reconstruction N={N}", flush=True)
    trj, U, cm = radial(N, K, C, nt, nspk)
    rng = np.random.default_rng(N[0])
    data = (rng.normal(size=(C,) + trj.shape[:2]) + 1j * rng.normal(size=(C,) + trj.shape[:2])).astype(np.complex64)
    cml = [dev(torch, c) for c in cm]
    A = msr.NFFTNormalOp(N, dev(torch, trj), dev(torch, U), cmaps=cml)
    lam_ref, km_ref = toeplitz.toeplitz_kernel(N, trj, U)
    km = A.prod.A.kmask_indcs.cpu().numpy()
    assert np.array_equal(km, km_ref + 1)
    lam = A.prod.A.Λ.cpu().numpy()
    print(f"  kernel rel err {rel_err(lam, lam_ref):.2e}", flush=True)
    assert rel_err(lam, lam_ref) < KERNEL_TOL
    b = msr.calculate_backprojection(dev(torch, data), dev(torch, trj), cml, U=dev(torch, U),
                                     density_compensation="radial_2D")
    bref = toeplitz.backprojection(data, trj, cm, U=U, density_compensation="radial_2D")
    print(f"  backprojection rel err {rel_err(b.cpu().numpy(), bref):.2e}", flush=True)
    assert rel_err(b.cpu().numpy(), bref) < NUFFT_TOL
    bd = b.reshape(-1)
    x = msr.cg(A, bd, maxiter=10, reltol=0.0, mode="device", precision="f64")
    bv = b.cpu().numpy()
    Af = lambda v: toeplitz.apply_normal(N, lam.astype(np.complex128), km - 1,
                                         cm.astype(np.complex128), v.reshape(bv.shape)).reshape(-1)
    ref = ocg.cg(Af, bv.reshape(-1).astype(np.complex128), maxiter=10, reltol=0.0)
    e = rel_err(x.cpu().numpy(), ref)
    print(f"  CG f64 rel err {e:.2e}", flush=True)
    assert e < 1e-7


def test_batch_with_line_apply_operators(torch, msr):
    """Line-kernel operators run their own passes in a batch's launch
    sequence: each member alone and inside the mix gives the same bits, and
    the batch CG matches the single-operator CG."""
    rng = np.random.default_rng(9)
    ops = []
    for N, K, C in (((32, 32), 3, 3), ((72, 40), 3, 2), ((32, 32), 9, 2), ((12, 10, 9), 2, 2)):
        lam, km, cm, x = make_case(rng, N, K, C)
        ops.append((make_op(torch, msr, N, lam, km, cm), dev(torch, x).reshape(-1)))
    big = msr.NFFTNormalOpBatch([a for a, _ in ops])
    ys = [torch.empty_like(x) for _, x in ops]
    big.mul_(ys, [x for _, x in ops])
    for (A, x), y in zip(ops, ys):
        one = msr.NFFTNormalOpBatch([A])
        y1 = torch.empty_like(x)
        one.mul_([y1], [x])
        assert torch.equal(y1, y)
        assert rel_err(y.cpu().numpy(), (A @ x).cpu().numpy()) <= 1e-6
    xs = [torch.empty_like(x) for _, x in ops]
    big.cg_(xs, [x for _, x in ops], maxiter=6, reltol=0.0, initially_zero=True, mode="graph",
            precision="f64")
    for (A, x), xb in zip(ops, xs):
        xr = msr.cg(A, x, maxiter=6, reltol=0.0, mode="device", precision="f64")
        assert rel_err(xb.cpu().numpy(), xr.cpu().numpy()) <= 1e-6


@pytest.mark.parametrize("N", [(45, 30), (27, 50)])
def test_cartesian_operator_at_odd_lengths(torch, msr, N):
    """The Cartesian operator's FFTs run at the image size: odd 2,3,5-smooth
    lengths (45, 27, 75 ...) take the non-exact plans."""
    K, C, nt = 2, 2, 4
    rng = np.random.default_rng(N[0])
    trj = traj_cartesian(N[0], N[1], nt)
    mask = rng.uniform(size=trj.shape[:2]) < 0.5
    U = rng.normal(size=(nt, K)) + 1j * rng.normal(size=(nt, K))
    U = np.ascontiguousarray(np.linalg.svd(U, full_matrices=False)[0].T).astype(np.complex64)
    cm = (rng.normal(size=(C,) + N[::-1]) + 1j * rng.normal(size=(C,) + N[::-1])).astype(np.complex64)
    x = (rng.normal(size=(K,) + N[::-1]) + 1j * rng.normal(size=(K,) + N[::-1])).astype(np.complex64)
    A = msr.FFTNormalOp(N, dev(torch, trj), dev(torch, U), cmaps=[dev(torch, c) for c in cm],
                        sample_mask=dev(torch, mask))
    Lr = ocart.kernel_cartesian(N, trj, U, mask)
    y = (A @ dev(torch, x).reshape(-1)).cpu().numpy().reshape(x.shape)
    yref = ocart.apply_normal(N, Lr, cm.astype(np.complex128), x.astype(np.complex128))
    print(f"cartesian N={N}: apply rel err {rel_err(y, yref):.2e}", flush=True)
    assert rel_err(y, yref) < APPLY_TOL


def test_limits_raise_cleanly(torch, msr):
    rng = np.random.default_rng(3)
    # K = 65: above the supported coefficient count
    G = 16 * 16
    km = np.arange(0, G, 3, dtype=np.int64)
    P = 65 * 66 // 2
    lam = (rng.normal(size=(km.size, P)) + 0j).astype(np.complex64)
    with pytest.raises(NotImplementedError):
        make_op(torch, msr, (8, 8), lam, km, None)
    # 2N = 2 * 7 = 14 is not 2,3,5-smooth: the reference's assertion
    lam2, km2, _, _ = make_case(rng, (7, 8), 2, 1)
    with pytest.raises(AssertionError):
        make_op(torch, msr, (7, 8), lam2, km2, None)
    # 2N = 2 * 1250 = 2500: smooth, but longer than the line FFTs
    with pytest.raises(NotImplementedError):
        make_op(torch, msr, (1250, 4), lam2[:0], km2[:0], None)
