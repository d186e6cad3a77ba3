"""Measured launch configuration of the fused 2D apply (msr_normalop_tune): every
candidate reproduces the untuned apply bit for bit, the choice carries over to
later operators of the same shape and to batches, captured CG graphs follow
it, and operators without fused 2D passes report no choice."""
import math

import numpy as np
import pytest

from conftest import load_package, rel_err
from oracle import toeplitz
from test_gpu_apply import make_case

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


def make_op(torch, msr, N, lam, km, cm):
    cml = None if cm is None else [dev(torch, c) for c in cm]
    return msr.NFFTNormalOp(N, dev(torch, lam), dev(torch, km + 1), cmaps=cml)


def apply_mode(torch, A, x, precision):
    if precision in ("f32", "mixed"):
        y = torch.empty_like(x)
        A.A.mul_(y, x)
        return y
    x64 = x.to(torch.complex128)
    y = torch.empty_like(x64)
    A.A.mul_f64_(y, x64, compact=precision == "f64c")
    return y


@pytest.mark.parametrize("precision", ["f32", "f64", "f64c"])
def test_tuning_keeps_the_result_bit_for_bit(torch, msr, precision):
    rng = np.random.default_rng(31)
    N, K, C = (128, 96), 4, 6
    lam, km, cm, x = make_case(rng, N, K, C)
    A = make_op(torch, msr, N, lam, km, cm)
    xd = dev(torch, x).reshape(-1)
    before = apply_mode(torch, A, xd, precision)
    cg_before = msr.cg(A, xd, maxiter=5, reltol=0.0, mode="graph", precision=precision)
    info = A.A.tune_(precision, nrep=3)
    t = info["times"]
    print(f"{precision}: chosen {info['chosen']}, times "
          + " ".join("-" if math.isnan(v) else f"{v:.3f}" for v in t), flush=True)
    assert info["chosen"] is not None
    assert len(t) == 16 and all(math.isfinite(v) and v > 0 for v in t[:12])  # Gb 1, 2, 4 <= C
    assert all(math.isnan(v) for v in t[12:])                               # Gb 8 > C
    g, oa, oc = info["chosen"]
    assert t[4 * int(math.log2(g)) + 2 * oa + oc] == min(t)
    after = apply_mode(torch, A, xd, precision)
    assert torch.equal(before, after)
    # the captured graph is re-captured with the new launch shapes: same bits
    cg_after = msr.cg(A, xd, maxiter=5, reltol=0.0, mode="graph", precision=precision)
    assert torch.equal(cg_before, cg_after)
    assert A.A.tune_(precision, nrep=0)["chosen"] == info["chosen"]
    # the oracle still agrees (the tuned apply is the apply)
    ref = toeplitz.apply_normal(N, lam.astype(np.complex128), km, cm.astype(np.complex128),
                                x.astype(np.complex128))
    assert rel_err(after.cpu().numpy().reshape(x.shape), ref) < (2e-6 if precision == "f32" else 2e-8)


def test_choice_carries_over_and_batches_stay_consistent(torch, msr):
    rng = np.random.default_rng(32)
    N, K, C = (96, 64), 3, 4
    lam, km, cm, x = make_case(rng, N, K, C)
    A = make_op(torch, msr, N, lam, km, cm)
    other = make_op(torch, msr, (64, 64), *make_case(rng, (64, 64), 3, 2)[:3])
    xd = dev(torch, x).reshape(-1)
    xo = dev(torch, rng.normal(size=3 * 64 * 64) + 0j).to(torch.complex64)
    early = msr.NFFTNormalOpBatch([A, other])  # made before the tuning
    info = A.A.tune_("f32", nrep=2)
    B = make_op(torch, msr, N, lam, km, cm)     # same shape, same context
    assert B.A.tune_("f32", nrep=0)["chosen"] == info["chosen"]
    late = msr.NFFTNormalOpBatch([A, other])
    ys = {}
    for name, bt in (("early", early), ("late", late)):
        y = [torch.empty_like(xd), torch.empty_like(xo)]
        bt.mul_(y, [xd, xo])
        ys[name] = y
    single = A @ xd
    assert torch.equal(ys["early"][0], ys["late"][0])
    assert torch.equal(ys["late"][1], ys["early"][1])
    assert rel_err(ys["late"][0].cpu().numpy(), single.cpu().numpy()) <= 1e-6


def test_single_coil_and_operators_without_fused_passes(torch, msr):
    rng = np.random.default_rng(33)
    lam, km, cm, x = make_case(rng, (64, 32), 2, 1)
    A = make_op(torch, msr, (64, 32), lam, km, cm)
    t = A.A.tune_("f64", nrep=2)["times"]
    assert all(math.isfinite(v) for v in t[:4]) and all(math.isnan(v) for v in t[4:])
    lam3, km3, cm3, _ = make_case(rng, (16, 16, 16), 2, 2)
    assert make_op(torch, msr, (16, 16, 16), lam3, km3, cm3).A.tune_("f32")["chosen"] is None
    lam4, km4, cm4, _ = make_case(rng, (72, 40), 2, 2)   # line-kernel apply
    assert make_op(torch, msr, (72, 40), lam4, km4, cm4).A.tune_("f32")["chosen"] is None
    with pytest.raises(msr.ArgumentError):
        A.A.tune_("f16")


def test_tuning_file_round_trip(torch, msr, tmp_path):
    rng = np.random.default_rng(34)
    N, K, C = (80, 48), 2, 3
    lam, km, cm, x = make_case(rng, N, K, C)
    A = make_op(torch, msr, N, lam, km, cm)
    chosen = A.A.tune_("f64c", nrep=2)["chosen"]
    path = tmp_path / "tune.txt"
    msr.save_tuning(str(path))
    lines = path.read_text().splitlines()
    assert lines[0].startswith("msr-tune 1 gfx")
    row = [ln for ln in lines[1:] if ln.split()[:7] == ["80", "48", "2", "3", "0", "0", "2"]]
    assert len(row) == 1 and tuple(int(v) for v in row[0].split()[7:]) == chosen
    assert msr.load_tuning(str(path)) == len(lines) - 1
    assert make_op(torch, msr, N, lam, km, cm).A.tune_("f64c", nrep=0)["chosen"] == chosen
    # another device type, a malformed entry, a missing file
    other = tmp_path / "other.txt"
    other.write_text("msr-tune 1 gfx000 1\n" + "\n".join(lines[1:]) + "\n")
    with pytest.raises(msr.ArgumentError):
        msr.load_tuning(str(other))
    bad = tmp_path / "bad.txt"
    bad.write_text(lines[0] + "\n80 48 2 3 0 0 2 3 0 0\n")  # Gb = 3
    with pytest.raises(OSError):
        msr.load_tuning(str(bad))
    with pytest.raises(OSError):
        msr.load_tuning(str(tmp_path / "missing.txt"))
