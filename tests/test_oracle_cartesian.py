"""Pins the Cartesian oracle (oracle/cartesian.py) against independent
constructions: the trajectory and the mask forms of the kernel agree, the
kernel is Hermitian per cell, the backprojection is 1/V times the adjoint of
the reference tests' FFT data simulation (test/reconstruct_cart_trj.jl:40-56),
the normal operator equals the backprojection of that simulation, and CG on
them recovers the image as the reference test asserts (rtol 1e-3)."""
import numpy as np
import pytest

from conftest import rel_err
from oracle import cartesian as ocart
from oracle.cg import cg


def traj_cartesian(nx, ny, nt):
    """traj_cartesian(Nx, Ny, 1, Nt; T=Int) of src/Trajectories.jl:15-33 in
    the (Nt, Nsamp, d) layout: sample s = x + Nx*y holds (x+1, y+1)."""
    x, y = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 1), indexing="xy")
    k = np.stack([x.reshape(-1), y.reshape(-1)], axis=-1)
    return np.broadcast_to(k, (nt,) + k.shape).astype(np.int32).copy()


def simulate(x, U, cmaps):
    """Data of test/reconstruct_cart_trj.jl:45-56: per coil and frame the
    image U*x .* S_c, ifftshift, fft, fftshift; (C, Nt, Nx*Ny)."""
    K = x.shape[0]
    img = np.einsum("kt,kyx->tyx", U, x)
    out = []
    for cm in cmaps:
        f = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(img * cm[None], axes=(1, 2))), axes=(1, 2))
        out.append(f.reshape(f.shape[0], -1))
    return np.stack(out)


def case(seed=0, nx=12, ny=10, nt=6, K=3, C=3, frac=0.6):
    rng = np.random.default_rng(seed)
    U = rng.normal(size=(nt, K)) + 1j * rng.normal(size=(nt, K))
    U, _, _ = np.linalg.svd(U, full_matrices=False)
    U = np.ascontiguousarray(U.T)
    cm = rng.normal(size=(C, ny, nx)) + 1j * rng.normal(size=(C, ny, nx))
    x = rng.normal(size=(K, ny, nx)) + 1j * rng.normal(size=(K, ny, nx))
    mask = rng.uniform(size=(nt, nx * ny)) < frac
    return traj_cartesian(nx, ny, nt), U, cm, x, mask


def test_kernel_trajectory_and_mask_forms_agree_and_are_hermitian():
    trj, U, cm, x, mask = case()
    L1 = ocart.kernel_cartesian((12, 10), trj, U, mask)
    M = mask.reshape(6, 10, 12)  # (Nt, Ny, Nx)
    L2 = ocart.kernel_cartesian_mask(M, U)
    assert np.allclose(L1, L2, rtol=0, atol=1e-13)
    assert np.allclose(L1, np.conj(np.transpose(L1, (0, 2, 1))), atol=1e-14)
    # a cell sampled by no frame has a zero kernel and is off the kmask
    Mz = M.copy()
    Mz[:, 0, 0] = False  # k-space corner -> cell (6, 5) after ifftshift
    km = ocart.kmask_of(ocart.kernel_cartesian_mask(Mz, U))
    assert 5 * 12 + 6 not in km
    assert np.array_equal(km, np.flatnonzero(np.fft.ifftshift(Mz.any(axis=0)).reshape(-1)))
    # repeated samples add
    trj2 = np.concatenate([trj, trj[:, :5]], axis=1)
    mask2 = np.concatenate([mask, np.ones((6, 5), bool)], axis=1)
    L3 = ocart.kernel_cartesian((12, 10), trj2, U, mask2)
    extra = L3 - L1
    q = ocart.sample_cells(trj[0, :5], (12, 10))
    assert np.allclose(extra[q], np.einsum("jt,kt->jk", np.conj(U), U)[None], atol=1e-13)
    assert np.allclose(np.delete(extra, q, axis=0), 0)


def test_backprojection_is_scaled_adjoint_of_the_simulation():
    trj, U, cm, x, mask = case(1)
    rng = np.random.default_rng(2)
    y = rng.normal(size=(3, 6, 120)) + 1j * rng.normal(size=(3, 6, 120))
    y = y * mask[None]
    b = ocart.backprojection(y, trj, cm, U=U, sample_mask=mask)
    lhs = np.vdot(x.reshape(-1), b.reshape(-1))
    Ex = simulate(x, U, cm) * mask[None]
    rhs = np.vdot(Ex.reshape(-1), y.reshape(-1)) / 120
    assert abs(lhs - rhs) < 1e-12 * abs(rhs)
    # per-coil form stacks the uncombined images
    bc = ocart.backprojection(y, trj, (12, 10), U=U, sample_mask=mask)
    assert bc.shape == (3, 3, 10, 12)
    assert rel_err(np.sum(np.conj(cm)[:, None] * bc, axis=0), b) < 1e-14


def test_normal_operator_is_backprojection_of_simulation():
    trj, U, cm, x, mask = case(3)
    Lam = ocart.kernel_cartesian((12, 10), trj, U, mask)
    Ax = ocart.apply_normal((12, 10), Lam, cm, x)
    ref = ocart.backprojection(simulate(x, U, cm), trj, cm, U=U, sample_mask=mask)
    assert rel_err(Ax, ref) < 1e-13
    # alpha/beta
    x0 = np.ones_like(x)
    assert rel_err(ocart.apply_normal((12, 10), Lam, cm, x, 0.5, 2.0, x0), 0.5 * Ax + 2.0 * x0) < 1e-14


def test_reference_cartesian_recon_recovers_image():
    # test/reconstruct_cart_trj.jl at a small size: 80% random mask, 20 CG
    # iterations, xr ≈ x rtol 1e-3
    nx, nt, K, C = 16, 20, 3, 4
    rng = np.random.default_rng(4)
    U = rng.normal(size=(nt, K)) + 1j * rng.normal(size=(nt, K))
    U, _, _ = np.linalg.svd(U, full_matrices=False)
    U = np.ascontiguousarray(U.T)
    x = rng.normal(size=(K, nx, nx)) + 1j * rng.normal(size=(K, nx, nx))
    cm = np.ones((C, nx, nx), complex)
    ix = np.arange(nx)
    cm[1] *= np.exp(1j * np.pi * ix / nx)[None, :]
    cm[2] *= np.exp(-1j * np.pi * ix / nx)[:, None]
    cm[3] *= np.exp(2j * np.pi * ix / nx)[None, :]
    cm /= np.linalg.norm(cm, axis=0)[None]
    mask = rng.uniform(size=(nt, nx * nx)) < 0.8
    trj = traj_cartesian(nx, nx, nt)
    data = simulate(x, U, cm)
    Lam = ocart.kernel_cartesian((nx, nx), trj, U, mask)
    b = ocart.backprojection(data, trj, cm, U=U, sample_mask=mask)
    xr = cg(lambda v: ocart.apply_normal((nx, nx), Lam, cm, v.reshape(x.shape)).reshape(-1),
            b.reshape(-1), 20).reshape(x.shape)
    assert rel_err(xr, x) < 1e-3
