"""msr_cg_noalloc, the CG entry point that allocates nothing (its vectors live
in one static device array per device), called through the C ABI: its result
equals msr_cg's in device mode (the same kernels in the same order) on two
calls in a row, the second on a smaller operator, at every precision; sizes
above the array and an apply scratch not sized for the precision are refused
before any kernel runs."""
import ctypes

import numpy as np
import pytest

from conftest import load_package, rel_err
from test_gpu_apply import make_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def noalloc(msr, torch, A, b, maxiter, precision):
    op = A.A
    x = torch.empty_like(b)
    it = ctypes.c_int()
    res = ctypes.c_double()
    msr._lib.call("msr_cg_noalloc", op.h, msr._ptr(x), msr._ptr(b), int(maxiter), 0.0, 0.0, 1,
                  msr._prec(precision), ctypes.byref(it), ctypes.byref(res),
                  msr._stream_ptr(op.device))
    return x, it.value, res.value


def make_op(torch, msr, rng, N, K, C):
    lam, km, cm, x = make_case(rng, N, K, C)
    dev = torch.device("cuda:0")
    A = msr.NFFTNormalOp(N, torch.from_numpy(lam).to(dev), torch.from_numpy(km + 1).to(dev),
                         cmaps=[torch.from_numpy(c).to(dev) for c in cm])
    return A, torch.from_numpy(x).to(dev).reshape(-1)


@pytest.mark.parametrize("precision", ["f64", "mixed", "f32"])
def test_noalloc_cg_equals_cg_on_two_calls(torch, precision):
    msr = load_package()
    rng = np.random.default_rng(314)
    big = make_op(torch, msr, rng, (96, 80), 4, 6)
    small = make_op(torch, msr, rng, (32, 48), 3, 2)
    for (A, b), iters in ((big, 9), (small, 6)):
        print(f"noalloc CG {precision}: N={A.A.shape} K={A.A.Ncoeff} {iters} iterations",
              flush=True)
        # msr_cg first: it sizes the apply scratch for the precision
        ref, info = msr.cg_(torch.empty_like(b), A, b, maxiter=iters, reltol=0.0,
                            initially_zero=True, mode="device", precision=precision, log=True)
        x, it, res = noalloc(msr, torch, A, b, iters, precision)
        torch.cuda.synchronize()
        same = bool(torch.equal(x, ref))
        e = rel_err(x.cpu().numpy(), ref.cpu().numpy())
        print(f"  identical {same}, rel diff {e:.1e}, iterations {it} / {info['iters']}",
              flush=True)
        assert same and e == 0.0
        assert it == info["iters"] == iters
        assert res == info["residual"]


def test_noalloc_refuses_oversized_and_unsized(torch):
    msr = load_package()
    rng = np.random.default_rng(2718)
    dev = torch.device("cuda:0")
    # 64^3 x K = 5 at FP64 needs about 84 MB of CG state: above the array
    N = (64, 64, 64)
    K = 5
    G = 8 * N[0] * N[1] * N[2]
    km = np.arange(G, dtype=np.int64)
    lam = np.zeros((G, K * (K + 1) // 2), np.complex64)
    lam[:, 0] = 1.0
    A = msr.NFFTNormalOp(N, torch.from_numpy(lam).to(dev), torch.from_numpy(km + 1).to(dev))
    b = torch.ones(K * N[0] * N[1] * N[2], dtype=torch.complex64, device=dev)
    with pytest.raises(msr.ArgumentError, match="above"):
        noalloc(msr, torch, A, b, 2, "f64")
    # a fresh operator has FP32 apply scratch only: FP64 is refused, FP32 runs
    A2, b2 = make_op(torch, msr, rng, (32, 32), 2, 2)
    with pytest.raises(msr.ArgumentError, match="not sized"):
        noalloc(msr, torch, A2, b2, 3, "f64")
    x, it, _ = noalloc(msr, torch, A2, b2, 3, "f32")
    assert it == 3 and torch.isfinite(torch.view_as_real(x)).all()
