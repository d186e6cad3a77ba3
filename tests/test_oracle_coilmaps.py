"""Pins the oracle's coil-map calibration front end (oracle/coilmaps.py)
against brute force: Julia's nextprod and the calib_size default on known
values, the calibration region by its defining inequality, the coil-wise
reconstruction against CG on the explicit NUDFT normal matrix (and image
recovery from exact data), and the calibration k-space against the centred
DFT sum written out."""
import numpy as np
import pytest

from conftest import load_package, rel_err
from oracle import coilmaps
from oracle.cg import cg

synth = load_package().synth


def test_nextprod_and_default_calib_size():
    # brute force over a range
    for n in range(1, 400):
        m = coilmaps.nextprod235(n)
        r = m
        for p in (2, 3, 5):
            while r % p == 0:
                r //= p
        assert r == 1 and m >= n
        assert all(coilmaps.nextprod235(k) == m for k in range(n, m + 1))
    assert coilmaps.calib_size_default((256, 256)) == (32, 32)
    assert coilmaps.calib_size_default((256, 128)) == (32, 16)
    assert coilmaps.calib_size_default((100, 80)) == (36, 27)  # 100÷3 = 33 -> 36, 80÷3 = 26 -> 27
    assert coilmaps.calib_size_default((128, 128, 96)) == (32, 32, 24)
    with pytest.raises(ZeroDivisionError):
        coilmaps.calib_size_default((16, 16))


def test_calib_region_is_the_scaled_box():
    rng = np.random.default_rng(0)
    trj = rng.uniform(-0.5, 0.5, size=(5, 40, 2)).astype(np.float32)
    mask = rng.uniform(size=(5, 40)) < 0.8
    tc, mc = coilmaps.calib_select(trj, (64, 48), (16, 16), mask)
    s = np.array([4.0, 3.0], np.float32)
    assert np.array_equal(tc, trj * s)
    box = (np.abs(trj[..., 0]) < 0.125) & (np.abs(trj[..., 1]) < 0.5 / 3)
    # away from the boundary the float32 products decide exactly like the box
    safe = (np.abs(np.abs(trj[..., 0]) - 0.125) > 1e-6) & (np.abs(np.abs(trj[..., 1]) - 0.5 / 3) > 1e-6)
    assert np.array_equal(mc[safe], (box & mask)[safe])
    assert np.all(np.abs(tc[mc]) < 0.5)
    # without a sample mask every in-box sample is kept
    _, m2 = coilmaps.calib_select(trj, (64, 48), (16, 16))
    assert np.array_equal(m2 & mask, mc)


def _explicit_normal(kc, cs):
    ry, rx = np.meshgrid(np.arange(cs[1]) - cs[1] // 2, np.arange(cs[0]) - cs[0] // 2, indexing="ij")
    r = np.stack([rx.reshape(-1), ry.reshape(-1)], 1)
    return np.exp(-2j * np.pi * kc.astype(np.float64) @ r.T)


@pytest.mark.parametrize("niter", [5, 40])
def test_coilwise_matches_cg_on_the_explicit_normal_matrix(niter):
    N, nt, nspk, C, cs = (64, 64), 12, 3, 3, (16, 16)
    trj = synth.traj_2d_radial_goldenratio(2 * max(N), nspk, nt)
    tc, mc = coilmaps.calib_select(trj, N, cs)
    rng = np.random.default_rng(1)
    data = rng.normal(size=(C,) + trj.shape[:2]) + 1j * rng.normal(size=(C,) + trj.shape[:2])
    x = coilmaps.reconstruct_coilwise(data, tc, cs, sample_mask=mc, niter=niter)
    E = _explicit_normal(tc[mc], cs)
    EhE = E.conj().T @ E
    y = data[:, mc]
    ref = np.stack([cg(lambda v: EhE @ v, E.conj().T @ y[c], niter, reltol=0.0)
                    for c in range(C)]).reshape(x.shape)
    assert rel_err(x, ref) < 1e-5
    # a coil-index or conjugation slip would not survive this
    assert rel_err(x[::-1], ref) > 0.1


def test_coilwise_recovers_images_from_exact_data():
    N, nt, nspk, C, cs = (64, 64), 16, 3, 2, (16, 16)
    trj = synth.traj_2d_radial_goldenratio(2 * max(N), nspk, nt)
    tc, mc = coilmaps.calib_select(trj, N, cs)
    rng = np.random.default_rng(2)
    xt = rng.normal(size=(C, 16, 16)) + 1j * rng.normal(size=(C, 16, 16))
    E = _explicit_normal(tc[mc], cs)
    data = np.zeros((C,) + trj.shape[:2], dtype=np.complex128)
    for c in range(C):
        data[c][mc] = E @ xt[c].reshape(-1)
    x = coilmaps.reconstruct_coilwise(data, tc, cs, sample_mask=mc, niter=100)
    assert rel_err(x, xt) < 1e-8


@pytest.mark.parametrize("shape", [(6, 4), (5, 7), (3, 4, 5)])
def test_calib_kspace_is_the_centred_dft(shape):
    rng = np.random.default_rng(3)
    d = len(shape)
    x = rng.normal(size=(2,) + shape) + 1j * rng.normal(size=(2,) + shape)
    k = coilmaps.calib_kspace(x, d)
    # out[q] = sum_r x[r] exp(-2 pi i sum_a (q_a - n_a//2)(r_a + n_a//2) / n_a)
    grids = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    for q in [tuple(0 for _ in shape), tuple(n - 1 for n in shape), tuple(n // 2 for n in shape),
              tuple((a + 1) % n for a, n in enumerate(shape))]:
        ph = sum((q[a] - shape[a] // 2) * (grids[a] + shape[a] // 2) / shape[a] for a in range(d))
        ref = np.sum(x * np.exp(-2j * np.pi * ph)[None], axis=tuple(range(1, d + 1)))
        assert np.allclose(k[(slice(None),) + q], ref, rtol=1e-12, atol=1e-10)
