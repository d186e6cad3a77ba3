"""GPU parity of the Toeplitz normal-operator apply and the device CG against
the FP64 oracle (same complex64 inputs, promoted to complex128 for the oracle).

Tolerances: one apply in FP32 (FFTs with double-precision twiddles, FP32
accumulation) is held to 2e-6 relative L2 error; a CG solve to 1e-4, the bound
the task sets for every result."""
import math

import numpy as np
import pytest

from conftest import load_package, rel_err
from oracle import cg as ocg
from oracle import toeplitz

pytestmark = pytest.mark.gpu

APPLY_TOL = 2e-6
CG_TOL = 1e-4


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.fixture(scope="module")
def msr():
    return load_package()


def random_packed_kernel(rng, n_grid, K, real, frac=0.8):
    """Random packed kernel on a random sorted mask: Hermitian by construction
    (the operator only reads the packed upper triangle)."""
    G = int(np.prod(n_grid))
    km = np.sort(rng.choice(G, size=int(frac * G), replace=False)).astype(np.int64)
    P = K * (K + 1) // 2
    if real:
        lam = rng.normal(size=(km.size, P)).astype(np.float32)
    else:
        lam = (rng.normal(size=(km.size, P)) + 1j * rng.normal(size=(km.size, P))).astype(np.complex64)
    return lam, km


def make_case(rng, N, K, C, real=False, identity=False):
    n_grid = tuple(2 * v for v in N)
    lam, km = random_packed_kernel(rng, n_grid, K, real)
    cm = None if identity else load_package().synth.gaussian_coil_maps(N, C, seed=int(rng.integers(1000)))
    x = (rng.normal(size=(K,) + N[::-1]) + 1j * rng.normal(size=(K,) + N[::-1])).astype(np.complex64)
    return lam, km, cm, x


def gpu_op(torch, msr, N, lam, km, cm):
    dev = torch.device("cuda:0")
    cml = None if cm is None else [torch.from_numpy(c).to(dev) for c in cm]
    return msr.NFFTNormalOp(N, torch.from_numpy(lam).to(dev), torch.from_numpy(km + 1).to(dev),
                            cmaps=cml)


CASES = [
    ((16, 16), 3, 4, False, False),
    ((16, 16), 3, 4, True, False),
    ((32, 16), 2, 3, False, False),
    ((16, 32), 4, 2, False, True),
    ((48, 80), 5, 3, False, False),
    ((96, 48), 6, 2, True, False),
    ((64, 64), 4, 8, False, False),
    ((8, 8), 8, 2, False, False),
    ((8, 8), 1, 1, True, False),
    ((160, 32), 7, 2, False, False),
    ((192, 16), 2, 2, False, False),
]


@pytest.mark.parametrize("N,K,C,real,identity", CASES)
def test_apply_matches_oracle(torch, msr, N, K, C, real, identity):
    print(f"apply case N={N} K={K} C={C} real={real} identity={identity}", flush=True)
    rng = np.random.default_rng(hash((N, K, C, real, identity)) % 2 ** 32)
    lam, km, cm, x = make_case(rng, N, K, C, real, identity)
    A = gpu_op(torch, msr, N, lam, km, cm)
    xd = torch.from_numpy(x).cuda().reshape(-1)
    out = (A @ xd).reshape(x.shape).cpu().numpy()
    ref = toeplitz.apply_normal(N, lam.astype(np.complex128 if not real else np.float64), km,
                                None if cm is None else cm.astype(np.complex128),
                                x.astype(np.complex128))
    assert rel_err(out, ref) < APPLY_TOL
    # the plain per-coil order gives the same result
    y = torch.empty_like(xd)
    A.A.mul_(y, xd, 1.0, 0.0, percoil=True)
    assert rel_err(y.reshape(x.shape).cpu().numpy(), ref) < APPLY_TOL
    # packed kernel and kmask round-trip through the operator (op.prod!.A.Λ)
    assert np.array_equal(A.prod.A.kmask_indcs.cpu().numpy(), km + 1)
    assert np.array_equal(A.prod.A.Λ.cpu().numpy(), lam)


F64_CASES = [
    ((16, 16), 3, 4, False, False),
    ((16, 16), 3, 4, True, False),
    ((48, 80), 5, 3, False, False),
    ((16, 32), 4, 2, False, True),
    ((8, 8), 8, 2, False, False),
    ((160, 32), 7, 2, False, False),
    ((256, 256), 5, 6, False, False),
    ((16, 16, 16), 3, 3, False, False),
    ((16, 8, 24), 2, 2, True, False),
]
F64_APPLY_TOL = 1e-12


@pytest.mark.parametrize("N,K,C,real,identity", F64_CASES)
def test_apply_f64_matches_oracle(torch, msr, N, K, C, real, identity):
    """The ComplexF64 apply of the "f64" CG precision: FP64 FFTs, kernel
    multiply and intermediates with the ComplexF32 kernel and coil maps
    promoted exactly, so it agrees with the FP64 oracle on the same (FP32)
    kernel and maps to ~1e-15."""
    print(f"f64 apply case N={N} K={K} C={C} real={real} identity={identity}", flush=True)
    rng = np.random.default_rng(hash((N, K, C, real, identity, 64)) % 2 ** 32)
    lam, km, cm, x = make_case(rng, N, K, C, real, identity)
    A = gpu_op(torch, msr, N, lam, km, cm)
    x64 = x.astype(np.complex128) + 1e-3j * rng.normal(size=x.shape)  # not FP32-representable
    xd = torch.from_numpy(x64).cuda().reshape(-1)
    y = torch.empty_like(xd)
    A.A.mul_f64_(y, xd)
    ref = toeplitz.apply_normal(N, lam.astype(np.float64 if real else np.complex128), km,
                                None if cm is None else cm.astype(np.complex128), x64)
    e = rel_err(y.reshape(x.shape).cpu().numpy(), ref)
    print(f"f64 apply rel err {e:.3e}", flush=True)
    assert e < F64_APPLY_TOL
    # alpha / beta in FP64
    z = torch.from_numpy(rng.normal(size=x64.size) + 1j * rng.normal(size=x64.size)).cuda()
    z0 = z.clone()
    A.A.mul_f64_(z, xd, 0.5 - 2j, 1.5 + 0.25j)
    refz = (1.5 + 0.25j) * z0.cpu().numpy() + (0.5 - 2j) * ref.reshape(-1)
    assert rel_err(z.cpu().numpy(), refz) < F64_APPLY_TOL
    with pytest.raises(msr.ArgumentError):
        A.A.mul_f64_(y.to(torch.complex64), xd)


def test_apply_alpha_beta_and_nan(torch, msr):
    rng = np.random.default_rng(11)
    N = (32, 32)
    lam, km, cm, x = make_case(rng, N, 3, 3)
    A = gpu_op(torch, msr, N, lam, km, cm)
    xd = torch.from_numpy(x).cuda().reshape(-1)
    Ax = (A @ xd).cpu().numpy()
    z = (rng.normal(size=x.size) + 1j * rng.normal(size=x.size)).astype(np.complex64)
    for percoil in (False, True):
        zd = torch.from_numpy(z).cuda()
        A.A.mul_(zd, xd, 0.5 - 2j, 1.5 + 0.25j, percoil=percoil)
        ref = (1.5 + 0.25j) * z.astype(np.complex128) + (0.5 - 2j) * Ax
        assert rel_err(zd.cpu().numpy(), ref) < APPLY_TOL
        nan = torch.full_like(xd, float("nan"))
        A.A.mul_(nan, xd, 1.0, 0.0, percoil=percoil)  # beta == 0 ignores NaN in x
        assert rel_err(nan.cpu().numpy(), Ax) < APPLY_TOL


def test_apply_full_size_256_20coils_k5(torch, msr):
    """The benchmark shape: 256^2 image, 512^2 grid, 20 coils, K = 5."""
    rng = np.random.default_rng(12)
    N = (256, 256)
    lam, km, cm, x = make_case(rng, N, 5, 20)
    A = gpu_op(torch, msr, N, lam, km, cm)
    xd = torch.from_numpy(x).cuda().reshape(-1)
    out = (A @ xd).reshape(x.shape).cpu().numpy()
    ref = toeplitz.apply_normal(N, lam.astype(np.complex128), km, cm.astype(np.complex128),
                                x.astype(np.complex128))
    assert rel_err(out, ref) < APPLY_TOL
    # linearity and self-adjointness at full size
    y = (rng.normal(size=x.shape) + 1j * rng.normal(size=x.shape)).astype(np.complex64)
    yd = torch.from_numpy(y).cuda().reshape(-1)
    Ay = (A @ yd).cpu().numpy()
    lhs = np.vdot(y.reshape(-1).astype(np.complex128), out.reshape(-1))
    rhs = np.vdot(Ay.astype(np.complex128), x.reshape(-1).astype(np.complex128))
    # the random kernel is Hermitian per voxel except on its diagonal; use the
    # oracle's value of the same inner products instead of exact symmetry
    ref_rhs = np.vdot(toeplitz.apply_normal(N, lam.astype(np.complex128), km,
                                            cm.astype(np.complex128),
                                            y.astype(np.complex128)).reshape(-1),
                      x.reshape(-1).astype(np.complex128))
    assert abs(rhs - ref_rhs) < 1e-5 * abs(ref_rhs)
    assert abs(lhs) > 0


def physical_case(N=(32, 32), K=3, C=4, nt=16, nspk=2, seed=0, real=False):
    synth = load_package().synth
    trj = synth.traj_2d_radial_goldenratio(2 * N[0], nspk, nt)
    U = synth.t2_basis(nt, K, seed=seed, complex_out=not real)
    cm = synth.gaussian_coil_maps(N, C, seed=seed)
    x = synth.ellipse_phantom(N, K, seed=seed)
    data = toeplitz.forward(x.astype(np.complex128), trj, U, cm.astype(np.complex128))
    b = toeplitz.backprojection(data, trj, cm.astype(np.complex128), U=U)
    lam, km = toeplitz.toeplitz_kernel(N, trj, U)
    lam = lam.astype(np.float32 if real else np.complex64)
    return trj, U, cm, b.astype(np.complex64), lam, km


# FP64 vectors and apply: only the FP64 rounding of the device arithmetic
# separates the solve from the oracle's (the result is returned in ComplexF32,
# whose rounding, ~3e-8, bounds the comparison)
CG_F64_TOL = 1e-7
# FP64 arithmetic with the FFT intermediates stored in 8 bytes (2^-27
# relative rounding per value, two roundings per apply): the applies of the
# solve differ from exact FP64 by ~1e-9, which 20 CG iterations on these
# indefinite operators amplify by up to ~1e3-1e4
CG_F64C_TOL = 2e-5
# one compact-storage apply against the FP64 oracle
APPLY_F64C_TOL = 2e-8

TOLS = {"f64": CG_F64_TOL, "f64c": CG_F64C_TOL}


@pytest.mark.parametrize("mode,precision", [("device", "f64"), ("graph", "f64"),
                                            ("device", "f64c"), ("graph", "f64c"),
                                            ("device", "mixed"), ("host", "mixed"),
                                            ("host", "f32"), ("graph", "f32")])
def test_cg_matches_fp64_oracle(torch, msr, mode, precision):
    print(f"cg mode={mode} precision={precision}", flush=True)
    N = (32, 32)
    trj, U, cm, b, lam, km = physical_case(N)
    A = gpu_op(torch, msr, N, lam, km, cm)
    bd = torch.from_numpy(b).cuda().reshape(-1)
    x, info = msr.cg_(torch.empty_like(bd), A, bd, maxiter=20, reltol=0.0, initially_zero=True,
                      mode=mode, precision=precision, log=True)
    assert info["iters"] == 20
    Af = lambda v: toeplitz.apply_normal(N, lam.astype(np.complex128), km, cm.astype(np.complex128),
                                         v.reshape(b.shape)).reshape(-1)
    ref = ocg.cg(Af, b.reshape(-1).astype(np.complex128), maxiter=20, reltol=0.0)
    e = rel_err(x.cpu().numpy(), ref)
    print(f"cg {mode}/{precision} rel err {e:.3e}", flush=True)
    assert e < TOLS.get(precision, CG_TOL)


def test_cg_graph_mode_first_use_of_a_precision(torch, msr):
    """A fresh operator whose first CG is graph-captured in a precision no
    earlier call used (FFT lengths 24 and 40 appear in no other test): every
    table and buffer of the apply is created before the capture begins."""
    rng = np.random.default_rng(21)
    N = (12, 20)
    lam, km, cm, x = make_case(rng, N, 2, 2)
    A = gpu_op(torch, msr, N, lam, km, cm)
    bd = torch.from_numpy(x).cuda().reshape(-1)
    Af = lambda v: toeplitz.apply_normal(N, lam.astype(np.complex128), km, cm.astype(np.complex128),
                                         v.reshape(x.shape)).reshape(-1)
    ref = ocg.cg(Af, x.reshape(-1).astype(np.complex128), maxiter=6, reltol=0.0)
    for precision in ("f64c", "f64", "mixed"):
        xg = msr.cg(A, bd, maxiter=6, reltol=0.0, mode="graph", precision=precision)
        e = rel_err(xg.cpu().numpy(), ref)
        print(f"graph first use {precision}: rel err {e:.3e}", flush=True)
        assert e < TOLS.get(precision, CG_TOL)


def test_cg_host_mode_rejects_f64_apply(torch, msr):
    N = (16, 16)
    trj, U, cm, b, lam, km = physical_case(N, K=2, C=2, nt=8, nspk=4)
    A = gpu_op(torch, msr, N, lam, km, cm)
    bd = torch.from_numpy(b).cuda().reshape(-1)
    with pytest.raises(NotImplementedError):
        msr.cg_(torch.empty_like(bd), A, bd, maxiter=2, mode="host", precision="f64")
    with pytest.raises(msr.ArgumentError):
        msr.cg_(torch.empty_like(bd), A, bd, maxiter=2, precision="f16")


def test_cg_default_tolerance_stops_like_reference(torch, msr):
    N = (16, 16)
    trj, U, cm, b, lam, km = physical_case(N, K=2, C=2, nt=8, nspk=4)
    A = gpu_op(torch, msr, N, lam, km, cm)
    bd = torch.from_numpy(b).cuda().reshape(-1)
    x, info = msr.cg_(torch.empty_like(bd), A, bd, maxiter=200, initially_zero=True, log=True)
    hist = []
    Af = lambda v: toeplitz.apply_normal(N, lam.astype(np.complex128), km, cm.astype(np.complex128),
                                         v.reshape(b.shape)).reshape(-1)
    ref = ocg.cg(Af, b.reshape(-1).astype(np.complex128), maxiter=200, history=hist)
    assert abs(info["iters"] - len(hist)) <= 1
    assert rel_err(x.cpu().numpy(), ref) < CG_TOL
    # graph replay gives the same result on every replay
    xs = []
    for _ in range(3):
        xg = torch.empty_like(bd)
        msr.cg_(xg, A, bd, maxiter=200, initially_zero=True, mode="graph")
        torch.cuda.synchronize()
        xs.append(xg.cpu().numpy())
    assert all(np.array_equal(xs[0], v) for v in xs[1:])
    assert rel_err(xs[0], ref) < CG_TOL


def test_cg_warm_start(torch, msr):
    N = (16, 16)
    trj, U, cm, b, lam, km = physical_case(N, K=2, C=2, nt=8, nspk=4, seed=3)
    A = gpu_op(torch, msr, N, lam, km, cm)
    bd = torch.from_numpy(b).cuda().reshape(-1)
    rng = np.random.default_rng(0)
    x0 = (rng.normal(size=b.size) + 1j * rng.normal(size=b.size)).astype(np.complex64)
    Af = lambda v: toeplitz.apply_normal(N, lam.astype(np.complex128), km, cm.astype(np.complex128),
                                         v.reshape(b.shape)).reshape(-1)
    ref = ocg.cg(Af, b.reshape(-1).astype(np.complex128), maxiter=10, reltol=0.0,
                 x0=x0.astype(np.complex128))
    for precision, tol in (("f64", CG_F64_TOL), ("f64c", CG_F64C_TOL), ("mixed", CG_TOL),
                           ("f32", CG_TOL)):
        xd = torch.from_numpy(x0).cuda()
        msr.cg_(xd, A, bd, maxiter=10, reltol=0.0, precision=precision)
        e = rel_err(xd.cpu().numpy(), ref)
        print(f"warm start {precision} rel err {e:.3e}", flush=True)
        assert e < tol


@pytest.mark.parametrize("N,K,C,real", [((16, 16), 3, 4, False), ((48, 80), 5, 3, False),
                                        ((96, 48), 6, 2, True), ((256, 256), 5, 20, False),
                                        ((160, 32), 7, 2, False), ((16, 16, 16), 3, 3, False),
                                        ((24, 20, 16), 2, 2, True), ((32, 32, 32), 4, 4, False)])
def test_compact_storage_apply_matches_oracle(torch, msr, N, K, C, real):
    """The FP64 apply with 8-byte intermediates (mul_f64_(compact=True), the
    apply of the "f64c" precision) against the FP64 oracle: the storage
    roundings of each value (2^-27 of the larger part; two per apply in 2D, four
    in 3D) bound the error."""
    print(f"compact apply N={N} K={K} C={C} real={real}", flush=True)
    rng = np.random.default_rng(int(np.prod(N)) + K)
    lam, km, cm, x = make_case(rng, N, K, C, real)
    A = gpu_op(torch, msr, N, lam, km, cm)
    xd = torch.from_numpy(x).cuda().reshape(-1).to(torch.complex128)
    y = torch.empty_like(xd)
    A.A.mul_f64_(y, xd, compact=True)
    ref = toeplitz.apply_normal(N, lam.astype(np.float64 if real else np.complex128), km,
                                cm.astype(np.complex128), x.astype(np.complex128))
    e = rel_err(y.cpu().numpy().reshape(x.shape), ref)
    y64 = torch.empty_like(xd)
    A.A.mul_f64_(y64, xd)
    e64 = rel_err(y64.cpu().numpy().reshape(x.shape), ref)
    print(f"compact apply rel err {e:.2e} (ComplexF64 intermediates {e64:.2e})", flush=True)
    assert e < APPLY_F64C_TOL
    assert e64 < 1e-12
    # alpha / beta on the same path
    z = torch.randn_like(xd)
    z0 = z.clone()
    A.A.mul_f64_(z, xd, alpha=0.5 - 1j, beta=2.0, compact=True)
    zr = 2.0 * z0.cpu().numpy() + (0.5 - 1j) * ref.reshape(-1)
    assert rel_err(z.cpu().numpy(), zr) < APPLY_F64C_TOL
