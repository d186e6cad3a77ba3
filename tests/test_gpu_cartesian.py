"""Cartesian normal operator and backprojection on the device (FFTNormalOp,
integer-trajectory calculate_backprojection) against the FP64 oracle
(oracle/cartesian.py): kernel and kmask from the trajectory, mask and dense
constructors, applies in ComplexF32 and ComplexF64, per-coil and combined
backprojections, the reference's Cartesian reconstruction test
(test/reconstruct_cart_trj.jl: CG recovers the image to rtol 1e-3), and the
operator inside the batch engine next to a non-Cartesian one."""
import numpy as np
import pytest

from conftest import load_package, rel_err
from oracle import cartesian as ocart
from test_oracle_cartesian import simulate, traj_cartesian

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


def basis(rng, nt, K, real=False):
    U = rng.normal(size=(nt, K)) + (0 if real else 1j * rng.normal(size=(nt, K)))
    U, _, _ = np.linalg.svd(U, full_matrices=False)
    return np.ascontiguousarray(U.T).astype(np.float32 if real else np.complex64)


def traj3d(n, nt):
    x, y, z = np.meshgrid(*[np.arange(1, m + 1) for m in n], indexing="ij")
    k = np.stack([x.T.reshape(-1), y.T.reshape(-1), z.T.reshape(-1)], -1)  # x fastest
    return np.broadcast_to(k, (nt,) + k.shape).astype(np.int32).copy()


CASES = [  # (N, K, C, nt, real, frac, three_d)
    ((32, 24), 3, 4, 6, False, 0.5, False),
    ((32, 32), 4, 3, 8, True, 0.3, False),
    ((16, 16, 24), 2, 2, 4, False, 0.4, True),
]


def make(N, K, C, nt, real, frac, three_d, seed=0):
    rng = np.random.default_rng(seed)
    trj = traj3d(N, nt) if three_d else traj_cartesian(N[0], N[1], nt)
    # a few repeated samples (several samples of one frame in one cell)
    trj = np.concatenate([trj, trj[:, :7]], axis=1)
    mask = rng.uniform(size=trj.shape[:2]) < frac
    U = basis(rng, nt, K, real)
    cm = (rng.normal(size=(C,) + N[::-1]) + 1j * rng.normal(size=(C,) + N[::-1])).astype(np.complex64)
    x = (rng.normal(size=(K,) + N[::-1]) + 1j * rng.normal(size=(K,) + N[::-1])).astype(np.complex64)
    return trj, mask, U, cm, x


@pytest.mark.parametrize("N,K,C,nt,real,frac,three_d", CASES)
def test_kernel_apply_and_constructors_match_oracle(torch, msr, N, K, C, nt, real, frac, three_d):
    trj, mask, U, cm, x = make(N, K, C, nt, real, frac, three_d)
    cml = [dev(torch, c) for c in cm]
    A = msr.FFTNormalOp(N, dev(torch, trj), dev(torch, U), cmaps=cml, sample_mask=dev(torch, mask))
    assert A.A.cartesian and A.A.P == K * K
    Lr = ocart.kernel_cartesian(N, trj, U, mask)
    km_ref = ocart.kmask_of(Lr)
    km = A.A.kmask_indcs.cpu().numpy()
    lam = A.A.Λ.cpu().numpy()                     # (Nkm, K, K): Julia (K, K, Nkm)
    assert np.array_equal(km, km_ref + 1)
    # lam[i, k, j] = Julia Λ[j, k, i] = Lr[q, j, k]
    ref = np.transpose(Lr[km_ref], (0, 2, 1))
    print(f"cartesian kernel {N} K={K}: rel err {rel_err(lam, ref):.2e}, nkm {len(km)}", flush=True)
    assert rel_err(lam, ref) < 1e-7
    assert np.array_equal(lam, np.conj(np.transpose(lam, (0, 2, 1))))  # exactly Hermitian
    # applies
    xd = dev(torch, x).reshape(-1)
    y = (A @ xd).cpu().numpy().reshape(x.shape)
    yref = ocart.apply_normal(N, Lr, cm.astype(np.complex128), x.astype(np.complex128))
    print(f"  apply f32 rel err {rel_err(y, yref):.2e}", flush=True)
    assert rel_err(y, yref) < 2e-6
    y64 = torch.empty(x.size, dtype=torch.complex128, device="cuda")
    A.A.mul_f64_(y64, xd.to(torch.complex128), 0.5, 0.0)
    # against the oracle on the device's (ComplexF32-rounded) kernel: FP64 arithmetic
    L32 = np.zeros_like(Lr)
    L32[km_ref] = np.transpose(lam, (0, 2, 1))
    yref32 = ocart.apply_normal(N, L32, cm.astype(np.complex128), x.astype(np.complex128))
    e64 = rel_err(y64.cpu().numpy().reshape(x.shape), 0.5 * yref32)
    print(f"  apply f64 rel err {e64:.2e}", flush=True)
    assert e64 < 1e-12
    # mask constructor (one sample per frame and cell): the same kernel, bit for bit
    if not three_d:
        trj0, mask0 = trj[:, :-7], mask[:, :-7]
        B = msr.FFTNormalOp(N, dev(torch, trj0), dev(torch, U), cmaps=cml,
                            sample_mask=dev(torch, mask0))
        M = mask0.reshape((nt,) + N[::-1])
        Bm = msr.FFTNormalOp(dev(torch, M), dev(torch, U), cmaps=cml)
        assert np.array_equal(Bm.A.kmask_indcs.cpu().numpy(), B.A.kmask_indcs.cpu().numpy())
        assert np.array_equal(Bm.A.Λ.cpu().numpy(), B.A.Λ.cpu().numpy())
        Lm = ocart.kernel_cartesian_mask(M, U)
        assert rel_err(Bm.A.Λ.cpu().numpy(), np.transpose(Lm[ocart.kmask_of(Lm)], (0, 2, 1))) < 1e-7
    # dense constructor from the full kernel: the same applies
    dense = np.zeros((int(np.prod(N)), K, K), dtype=lam.dtype)
    dense[km - 1] = lam                             # [q, k, j] = Julia Λ[j, k, q]
    D = msr.FFTNormalOp(dev(torch, dense.reshape(N[::-1] + (K, K))), cmaps=cml)
    assert np.array_equal(D.A.kmask_indcs.cpu().numpy(), km)
    assert np.array_equal((D @ xd).cpu().numpy(), (A @ xd).cpu().numpy())


# U = I(Nt) with 11 frames gives K = 11: the gridding's coefficient blocks of 8
# run once whole and once partial
@pytest.mark.parametrize("per_coil,use_U,three_d,nt", [(False, True, False, 6), (True, True, False, 6),
                                                       (False, False, False, 6), (False, True, True, 4),
                                                       (True, False, False, 11)])
def test_backprojection_matches_oracle(torch, msr, per_coil, use_U, three_d, nt):
    N = (16, 16, 24) if three_d else (32, 24)
    trj, mask, U, cm, x = make(N, 3, 3, nt, False, 0.6, three_d, seed=1)
    rng = np.random.default_rng(2)
    data = (rng.normal(size=(3,) + trj.shape[:2]) + 1j * rng.normal(size=(3,) + trj.shape[:2])).astype(np.complex64)
    arg = N if per_coil else [dev(torch, c) for c in cm]
    out = msr.calculate_backprojection(dev(torch, data), dev(torch, trj), arg,
                                       U=dev(torch, U) if use_U else None,
                                       sample_mask=dev(torch, mask)).cpu().numpy()
    ref = ocart.backprojection(data, trj, N if per_coil else cm, U=U if use_U else None,
                               sample_mask=mask)
    print(f"cartesian bp per_coil={per_coil} U={use_U} 3d={three_d} nt={nt}: rel err "
          f"{rel_err(out, ref):.2e}", flush=True)
    assert out.shape == ref.shape
    assert rel_err(out, ref) < 2e-6
    with pytest.raises(msr.ArgumentError):
        msr.calculate_backprojection(dev(torch, data), dev(torch, trj), arg, U=dev(torch, U),
                                     density_compensation="radial_2D")


def reference_cart_case(nx=64, nt=40, K=4, C=9, seed=3):
    """test/reconstruct_cart_trj.jl:12-57 with a random image: phase-ramp coil
    maps normalised over coils, SVD basis, FFT data, 80% random mask."""
    rng = np.random.default_rng(seed)
    x = (rng.normal(size=(K, nx, nx)) + 1j * rng.normal(size=(K, nx, nx))).astype(np.complex64)
    cm = np.ones((C, nx, nx), dtype=np.complex128)
    i = np.arange(1, nx + 1)
    for c, (ax, s) in enumerate([(1, 1), (1, -1), (0, 1), (0, -1), (1, 2), (1, -2), (0, 2), (0, -2)], 1):
        ph = np.exp(1j * np.pi * s * i / nx)
        cm[c] *= ph[None, :] if ax == 1 else ph[:, None]
    cm /= np.linalg.norm(cm, axis=0)[None]
    U = basis(rng, nt, K)
    data = simulate(x.astype(np.complex128), U.astype(np.complex128), cm).astype(np.complex64)
    mask = rng.uniform(size=(nt, nx * nx)) < 0.8
    return x, cm.astype(np.complex64), U, data, traj_cartesian(nx, nx, nt), mask


@pytest.mark.parametrize("mode,precision", [("graph", "f64"), ("host", "f32")])
def test_reference_cartesian_reconstruction(torch, msr, mode, precision):
    x, cm, U, data, trj, mask = reference_cart_case()
    cml = [dev(torch, c) for c in cm]
    A = msr.FFTNormalOp((64, 64), dev(torch, trj), dev(torch, U), cmaps=cml,
                        sample_mask=dev(torch, mask))
    # Λ[:, :, i]' ≈ Λ[:, :, i]  (test/reconstruct_cart_trj.jl:67-73)
    lam = A.A.Λ.cpu().numpy()
    assert np.allclose(lam, np.conj(np.transpose(lam, (0, 2, 1))))
    b = msr.calculate_backprojection(dev(torch, data), dev(torch, trj), cml, U=dev(torch, U),
                                     sample_mask=dev(torch, mask))
    xr = msr.cg(A, b.reshape(-1), maxiter=20, mode=mode, precision=precision).cpu().numpy()
    err = rel_err(xr.reshape(x.shape), x)
    print(f"cartesian recon {mode}/{precision}: rel err vs truth {err:.2e}", flush=True)
    assert err < 1e-3


def test_cartesian_operator_in_a_batch_and_errors(torch, msr):
    trj, mask, U, cm, x = make((32, 24), 3, 2, 6, False, 0.5, False, seed=4)
    cml = [dev(torch, c) for c in cm]
    A = msr.FFTNormalOp((32, 24), dev(torch, trj), dev(torch, U), cmaps=cml,
                        sample_mask=dev(torch, mask))
    synth = msr.synth
    t2 = synth.traj_2d_radial_goldenratio(64, 12, 6)  # well sampled: CG is not chaotic
    B = msr.NFFTNormalOp((32, 24), dev(torch, t2), dev(torch, U), cmaps=cml)
    bA = dev(torch, x).reshape(-1)
    bB = dev(torch, x[::-1].copy()).reshape(-1)
    ra = msr.cg(A, bA, maxiter=6, reltol=0.0, mode="device").cpu().numpy()
    rb = msr.cg(B, bB, maxiter=6, reltol=0.0, mode="device").cpu().numpy()
    batch = msr.NFFTNormalOpBatch([A, B])
    xs = [torch.empty_like(bA), torch.empty_like(bB)]
    batch.cg_(xs, [bA, bB], maxiter=6, reltol=0.0, initially_zero=True, mode="graph")
    # the batch's vector reductions are blocked differently from msr_cg's
    assert rel_err(xs[0].cpu().numpy(), ra) <= 1e-6
    assert rel_err(xs[1].cpu().numpy(), rb) <= 1e-6
    one = msr.NFFTNormalOpBatch([A])
    x1 = [torch.empty_like(bA)]
    one.cg_(x1, [bA], maxiter=6, reltol=0.0, initially_zero=True, mode="graph")
    assert torch.equal(x1[0], xs[0])  # alone and beside another operator: bit-identical
    # a shared-kernel Cartesian operator applies like the original
    S = msr.NFFTNormalOp_shared(A, cmaps=cml)
    assert S.A.cartesian
    assert np.array_equal((S @ bA).cpu().numpy(), (A @ bA).cpu().numpy())
    with pytest.raises(msr.ArgumentError):  # NFFTNormalOp on integer indices
        msr.NFFTNormalOp((32, 24), dev(torch, trj), dev(torch, U))
    with pytest.raises(msr.ArgumentError):  # FFTNormalOp on a float trajectory
        msr.FFTNormalOp((32, 24), dev(torch, t2), dev(torch, U))
    with pytest.raises(NotImplementedError):
        A.A.mul_(torch.empty_like(bA), bA, percoil=True)
    with pytest.raises(NotImplementedError):
        msr.FFTNormalOp((31, 24), dev(torch, trj), dev(torch, U))  # no FFT of prime length 31


def test_cartesian_backprojection_inverts_nufft_on_the_grid(torch, msr):
    """test/backprojection_cart.jl: data simulated by the (type-2 NUFFT)
    forward model at the float Cartesian grid k = (i - 1 - N/2)/N, then
    backprojected with the equivalent integer trajectory i, returns the image
    (coil maps normalised over coils; the reference asserts atol 4e-5)."""
    Nx, C = 32, 9
    rng = np.random.default_rng(6)
    x = (rng.normal(size=(1, Nx, Nx)) + 1j * rng.normal(size=(1, Nx, Nx))).astype(np.complex64)
    cm = msr.synth.gaussian_coil_maps((Nx, Nx), C, seed=6).astype(np.complex128)
    cm /= np.linalg.norm(cm, axis=0)[None]
    cm = cm.astype(np.complex64)
    cml = [dev(torch, c) for c in cm]
    ti = traj_cartesian(Nx, Nx, 1)                            # integer 1..N
    tf = ((ti - 1 - Nx // 2) / Nx).astype(np.float32)         # traj_cartesian(...; T=Float32)
    U = np.ones((1, 1), dtype=np.complex64)
    data = msr.forward(dev(torch, x), dev(torch, tf), cmaps=cml, U=dev(torch, U))
    reco = msr.calculate_backprojection(data, dev(torch, ti), cml, U=dev(torch, U)).cpu().numpy()
    err = np.abs(reco - x).max()
    print(f"cartesian backprojection of NUFFT grid data: max abs err {err:.2e}", flush=True)
    assert err < 4e-5
