"""Batched applies and lockstep CG (msr_batch_*): a mix of operators of uneven
image size, coil count, K, real/complex kernel and dimension gives, for every
operator, the result of its own single-operator call, and the FP64 oracle's
within the single-operator tolerances (apply 2e-6, CG 1e-4).  Bit-identity is
asserted between an operator alone in a batch and the same operator inside a
larger batch (same kernels, same block partition); the single-operator path is
held to 1e-6 against the batched one."""
import numpy as np
import pytest

from conftest import load_package, rel_err
from oracle import cg as ocg
from oracle import toeplitz
from test_gpu_apply import make_case

pytestmark = pytest.mark.gpu

MIX = [  # (N, K, C, real, identity)
    ((32, 32), 3, 4, False, False),
    ((48, 32), 5, 3, False, False),
    ((32, 32), 3, 6, True, False),
    ((64, 40), 4, 2, False, True),
    ((16, 16, 16), 2, 3, False, False),
    ((32, 32), 6, 5, False, False),
]


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.fixture(scope="module")
def mix(torch):
    msr = load_package()
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(77)
    jobs = []
    for N, K, C, real, ident in MIX:
        lam, km, cm, x = make_case(rng, N, K, C, real, ident)
        cml = None if cm is None else [torch.from_numpy(c).to(dev) for c in cm]
        A = msr.NFFTNormalOp(N, torch.from_numpy(lam).to(dev), torch.from_numpy(km + 1).to(dev),
                             cmaps=cml)
        jobs.append(dict(N=N, K=K, lam=lam, km=km, cm=cm, x=x, A=A,
                         xd=torch.from_numpy(x).to(dev).reshape(-1)))
    return jobs


def oracle_apply(j):
    lam = j["lam"].astype(np.float64 if j["lam"].dtype == np.float32 else np.complex128)
    cm = None if j["cm"] is None else j["cm"].astype(np.complex128)
    return toeplitz.apply_normal(j["N"], lam, j["km"], cm, j["x"].astype(np.complex128))


def test_batch_apply_matches_single_and_oracle(torch, mix):
    msr = load_package()
    print(f"batch apply of {len(mix)} operators", flush=True)
    B = msr.NFFTNormalOpBatch([j["A"] for j in mix])
    ys = [torch.empty_like(j["xd"]) for j in mix]
    B.mul_(ys, [j["xd"] for j in mix])
    for j, y in zip(mix, ys):
        single = (j["A"] @ j["xd"]).cpu().numpy()
        yb = y.cpu().numpy()
        e_single = rel_err(yb, single)
        e_ref = rel_err(yb.reshape(j["x"].shape), oracle_apply(j))
        print(f"  N={j['N']} K={j['K']}: vs single {e_single:.2e} (bit-exact "
              f"{np.array_equal(yb, single)}), vs oracle {e_ref:.2e}", flush=True)
        assert e_single <= 1e-6
        assert e_ref < 2e-6
    # alpha / beta: y = 0.5 y - 2 A x
    y0 = [y.clone() for y in ys]
    B.mul_(ys, [j["xd"] for j in mix], alpha=-2.0, beta=0.5)
    for a, b in zip(ys, y0):
        assert rel_err(a.cpu().numpy(), (-1.5 * b).cpu().numpy()) < 1e-6


def test_batch_member_is_bit_identical_to_batch_of_one(torch, mix):
    msr = load_package()
    big = msr.NFFTNormalOpBatch([j["A"] for j in mix])
    yb = [torch.empty_like(j["xd"]) for j in mix]
    big.mul_(yb, [j["xd"] for j in mix])
    for i, j in enumerate(mix):
        one = msr.NFFTNormalOpBatch([j["A"]])
        y1 = torch.empty_like(j["xd"])
        one.mul_([y1], [j["xd"]])
        assert torch.equal(y1, yb[i]), f"operator {i} differs alone vs in the batch"


@pytest.mark.parametrize("mode,precision", [("device", "f64"), ("graph", "f64"),
                                            ("graph", "f64c")])
def test_batch_cg_matches_single_cg(torch, mix, mode, precision):
    msr = load_package()
    B = msr.NFFTNormalOpBatch([j["A"] for j in mix])
    bs = [j["xd"] for j in mix]
    xs = [torch.empty_like(b) for b in bs]
    for rep in range(2):  # graph: capture, then replay
        B.cg_(xs, bs, maxiter=12, reltol=0.0, initially_zero=True, mode=mode,
              precision=precision)
    for j, x in zip(mix, xs):
        xr = msr.cg(j["A"], j["xd"], maxiter=12, reltol=0.0, mode="device", precision=precision)
        e = rel_err(x.cpu().numpy(), xr.cpu().numpy())
        print(f"  {mode}/{precision} N={j['N']}: batch vs single CG {e:.2e} (bit-exact "
              f"{torch.equal(x, xr)})", flush=True)
        assert e <= 1e-6


def test_batch_cg_vs_oracle_and_early_stop(torch, mix):
    """Default tolerance: operators stop at their own iteration counts; each
    solution matches the FP64 oracle run for the same number of iterations."""
    msr = load_package()
    sub = [mix[0], mix[3], mix[4]]
    B = msr.NFFTNormalOpBatch([j["A"] for j in sub])
    bs = [j["xd"] for j in sub]
    xs = [torch.empty_like(b) for b in bs]
    _, info = B.cg_(xs, bs, maxiter=40, reltol=1e-3, initially_zero=True, mode="device", log=True)
    print(f"  iterations {info['iters']}", flush=True)
    for j, x, it in zip(sub, xs, info["iters"]):
        lam = j["lam"].astype(np.float64 if j["lam"].dtype == np.float32 else np.complex128)
        cm = None if j["cm"] is None else j["cm"].astype(np.complex128)
        op = toeplitz.NormalOp(j["N"], lam, j["km"], cm)
        ref = ocg.cg(op, j["x"].astype(np.complex128).reshape(-1), maxiter=it, reltol=0.0)
        e = rel_err(x.cpu().numpy(), ref)
        print(f"  N={j['N']}: {it} it, vs FP64 {e:.2e}", flush=True)
        assert e < 1e-4


def test_continuous_batching_staggered_joins(torch, mix):
    """Jobs join a running batch between iterations and leave when done: each
    result is bit-identical to the same job solved alone with the same calls."""
    msr = load_package()
    A, Bj, Cj = mix[0], mix[1], mix[4]
    outs = {k: torch.empty_like(j["xd"]) for k, j in (("a", A), ("b", Bj), ("c", Cj))}
    st = torch.cuda.current_stream()
    # a starts alone, b joins after 4 iterations, c after 7; 10 iterations each
    ba = msr.NFFTNormalOpBatch([A["A"]])
    ba.cg_start([outs["a"]], [A["xd"]], maxiter=10, reltol=0.0)
    ba.cg_iterate(4)
    bb = msr.NFFTNormalOpBatch([Bj["A"]])
    bb.cg_start([outs["b"]], [Bj["xd"]], maxiter=10, reltol=0.0)
    ab = msr.NFFTNormalOpBatch([A["A"], Bj["A"]])
    ab.cg_iterate(3)
    bc = msr.NFFTNormalOpBatch([Cj["A"]])
    bc.cg_start([outs["c"]], [Cj["xd"]], maxiter=10, reltol=0.0)
    abc = msr.NFFTNormalOpBatch([A["A"], Bj["A"], Cj["A"]])
    abc.cg_iterate(3)            # a reaches 10 here
    its, _, done = abc.cg_status()
    assert its == [10, 6, 3] and done == [True, False, False]
    ba.cg_finish()
    bcb = msr.NFFTNormalOpBatch([Bj["A"], Cj["A"]])
    bcb.cg_iterate(4)            # b reaches 10
    bb.cg_finish()
    bc.cg_iterate(3)             # c reaches 10
    bc.cg_finish()
    st.synchronize()
    for k, j in (("a", A), ("b", Bj), ("c", Cj)):
        alone = torch.empty_like(j["xd"])
        s = msr.NFFTNormalOpBatch([j["A"]])
        s.cg_start([alone], [j["xd"]], maxiter=10, reltol=0.0)
        s.cg_iterate(10)
        s.cg_finish()
        assert torch.equal(alone, outs[k]), f"job {k} differs alone vs in the mix"


def test_batch_errors(torch, mix):
    msr = load_package()
    with pytest.raises(msr.ArgumentError):
        msr.NFFTNormalOpBatch([mix[0]["A"], mix[0]["A"]])
    B = msr.NFFTNormalOpBatch([mix[0]["A"], mix[1]["A"]])
    with pytest.raises(msr.ArgumentError):
        B.mul_([mix[0]["xd"]], [mix[0]["xd"]])
    with pytest.raises(msr.ArgumentError):
        B.mul_([mix[1]["xd"], mix[0]["xd"]], [mix[0]["xd"], mix[1]["xd"]])
    with pytest.raises(NotImplementedError):
        B.cg_([torch.empty_like(mix[0]["xd"]), torch.empty_like(mix[1]["xd"])],
              [mix[0]["xd"], mix[1]["xd"]], maxiter=2, mode="host")
