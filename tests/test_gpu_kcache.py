"""Kernel persistence and cache (msr_kernel_*).  A saved kernel file holds
exactly the operator's packed kernel and kmask in the documented layout (read
back by an independent numpy reader); an operator loaded from it, or from a
file the numpy writer made from the same arrays, applies and solves
bit-identically to the original.  The protocol key changes with any input the
kernel depends on and with nothing else.  The cache builds once per protocol,
then serves from device memory or its directory, evicts least recently used
kernels beyond its budget, and rebuilds a corrupt file."""
import os

import numpy as np
import pytest

import kernel_file as kf
from conftest import load_package
from test_gpu_3d import kooshball_case
from test_gpu_nufft import radial_case

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


def applies(torch, A, B, x):
    xd = dev(torch, x).reshape(-1)
    return (A @ xd).cpu().numpy(), (B @ xd).cpu().numpy()


CASES = [  # (N, K, C, nt, nspk, real, three_d)
    ((24, 40), 4, 3, 5, 4, False, False),
    ((32, 32), 3, 2, 6, 4, True, False),
    ((16, 16, 16), 2, 3, 8, 6, False, True),
]


@pytest.mark.parametrize("N,K,C,nt,nspk,real,three_d", CASES)
def test_save_load_round_trip(torch, msr, tmp_path, N, K, C, nt, nspk, real, three_d):
    if three_d:
        trj, U, cm, x = kooshball_case(N, K, C, nt, nspk)
    else:
        trj, U, cm, x = radial_case(N, K, C, nt, nspk, real=real)
    cml = [dev(torch, c) for c in cm]
    trj_d, U_d = dev(torch, trj), dev(torch, U)
    A = msr.NFFTNormalOp(N, trj_d, U_d, cmaps=cml)
    key = msr.kernel_key(N, trj_d, U_d)
    path = str(tmp_path / "a.msrk")
    msr.save_kernel(A, path, key=key)
    lam = A.prod.A.Λ.cpu().numpy()
    km = A.prod.A.kmask_indcs.cpu().numpy()
    # the file holds the documented layout
    info, flam, fkm = kf.read(path)
    assert info == msr.kernel_file_info(path)
    assert info["img_shape"] == tuple(N) and info["Ncoeff"] == K and info["key"] == key
    assert info["lambda_real"] == real and info["nkm"] == len(km)
    assert os.path.getsize(path) == 128 + len(km) * (8 + lam.shape[1] * (4 if real else 8))
    assert np.array_equal(flam, lam) and np.array_equal(fkm, km)
    # loaded operator: same kernel, same applies and solves, bit for bit
    B = msr.load_kernel(path, cmaps=cml, key=key)
    assert np.array_equal(B.prod.A.Λ.cpu().numpy(), lam)
    assert np.array_equal(B.prod.A.kmask_indcs.cpu().numpy(), km)
    ya, yb = applies(torch, A, B, x)
    assert np.array_equal(ya, yb)
    b = dev(torch, x).reshape(-1)
    xa = msr.cg(A, b, maxiter=6, reltol=0.0, mode="device").cpu().numpy()
    xb = msr.cg(B, b, maxiter=6, reltol=0.0, mode="device").cpu().numpy()
    assert np.array_equal(xa, xb)
    # a file made by the numpy writer from the same arrays loads to the same operator
    p2 = str(tmp_path / "w.msrk")
    kf.write(p2, N, lam, km)
    C2 = msr.load_kernel(p2, cmaps=cml)
    assert np.array_equal(applies(torch, A, C2, x)[1], ya)
    # a key of another protocol is refused
    with pytest.raises(msr.ArgumentError):
        msr.load_kernel(path, cmaps=cml, key="1" * 32)


def test_damaged_kernel_file_is_refused_before_any_device_work(torch, msr, tmp_path):
    N = (24, 24)
    trj, U, cm, x = radial_case(N, 2, 1, 4, 3)
    A = msr.NFFTNormalOp(N, dev(torch, trj), dev(torch, U))
    path = str(tmp_path / "a.msrk")
    msr.save_kernel(A, path)
    raw = bytearray(open(path, "rb").read())
    raw[-3] ^= 0x40
    open(path, "wb").write(bytes(raw))
    with pytest.raises(OSError, match="checksum"):
        msr.load_kernel(path)
    with pytest.raises(OSError):
        msr.load_kernel(str(tmp_path / "none.msrk"))


def test_protocol_key_follows_the_kernel_inputs(torch, msr):
    N, K, nt, nspk = (32, 32), 3, 6, 4
    trj, U, cm, x = radial_case(N, K, 1, nt, nspk)
    k0 = msr.kernel_key(N, dev(torch, trj), dev(torch, U))
    assert len(k0) == 32
    # deterministic, and a function of values, not buffers
    assert msr.kernel_key(N, dev(torch, trj.copy()), dev(torch, U.copy())) == k0
    assert msr.kernel_key(N, dev(torch, trj), dev(torch, U)) == k0
    # an all-true mask selects the same samples: same kernel, same key
    mask = np.ones(trj.shape[:2], dtype=bool)
    assert msr.kernel_key(N, dev(torch, trj), dev(torch, U), sample_mask=dev(torch, mask)) == k0
    keys = {k0}

    def new(k):
        assert k not in keys
        keys.add(k)

    t2 = trj.copy()
    t2[3, 5, 1] = np.nextafter(t2[3, 5, 1], np.float32(1))  # one ulp of one coordinate
    new(msr.kernel_key(N, dev(torch, t2), dev(torch, U)))
    U2 = U.copy()
    U2[1, 2] = U2[1, 2] + np.float32(1e-6) * abs(U2[1, 2])
    new(msr.kernel_key(N, dev(torch, trj), dev(torch, U2)))
    mask[2, 7] = False  # one sample fewer
    new(msr.kernel_key(N, dev(torch, trj), dev(torch, U), sample_mask=dev(torch, mask)))
    new(msr.kernel_key((32, 30), dev(torch, trj), dev(torch, U)))
    new(msr.kernel_key(N, dev(torch, trj), dev(torch, U[:2].copy())))
    Ur = np.ascontiguousarray(U.real).astype(np.float32)  # real basis vs the same values complex
    kr = msr.kernel_key(N, dev(torch, trj), dev(torch, Ur))
    kc = msr.kernel_key(N, dev(torch, trj), dev(torch, Ur.astype(np.complex64)))
    assert kr != kc
    new(kr)
    if np.array_equal(Ur.astype(np.complex64), U):  # the basis was real-valued already
        assert kc == k0
    else:
        new(kc)


def test_cache_builds_once_serves_from_memory_and_disk_and_evicts(torch, msr, tmp_path):
    N1, N2 = (32, 32), (24, 40)
    t1, U1, cm1, x1 = radial_case(N1, 3, 2, 6, 4, seed=1)
    t2, U2, cm2, x2 = radial_case(N2, 4, 3, 5, 4, seed=2)
    c1 = [dev(torch, c) for c in cm1]
    c1b = [dev(torch, np.conj(c)) for c in cm1]  # another subject, same protocol
    c2 = [dev(torch, c) for c in cm2]
    p1 = (N1, dev(torch, t1), dev(torch, U1))
    p2 = (N2, dev(torch, t2), dev(torch, U2))
    cache = msr.KernelCache(directory=str(tmp_path))
    A = cache.get(*p1, cmaps=c1)
    assert cache.last_source == "built"
    k1 = cache.last_key
    assert k1 == msr.kernel_key(*p1)
    f1 = tmp_path / f"{k1}.msrk"
    assert f1.exists() and kf.read(str(f1))[0]["key"] == k1
    B = cache.get(*p1, cmaps=c1b)
    assert cache.last_source == "memory"
    fresh = msr.NFFTNormalOp(N1, p1[1], p1[2], cmaps=c1b)
    assert np.array_equal(*applies(torch, B, fresh, x1))
    assert msr.footprint(B)[1] == msr.footprint(A)[1]
    # a new cache on the same directory reads the file
    other = msr.KernelCache(directory=str(tmp_path))
    D = other.get(*p1, cmaps=c1b)
    assert other.last_source == "disk"
    assert np.array_equal(*applies(torch, D, fresh, x1))
    st = cache.stats()
    assert st["entries"] == 1 and st["built"] == 1 and st["memory"] == 1 and st["disk"] == 0
    assert st["bytes"] == msr.footprint(A)[1]

    # eviction: a budget that holds one of the two kernels
    kb1 = msr.footprint(A)[1]
    E = msr.NFFTNormalOp(N2, p2[1], p2[2], cmaps=c2)
    kb2 = msr.footprint(E)[1]
    small = msr.KernelCache(budget_bytes=max(kb1, kb2) + min(kb1, kb2) // 2)
    a1 = small.get(*p1, cmaps=c1)
    y_before = (a1 @ dev(torch, x1).reshape(-1)).cpu().numpy()
    assert small.last_source == "built"
    small.get(*p1, cmaps=c1)
    assert small.last_source == "memory"
    e2 = small.get(*p2, cmaps=c2)
    assert small.last_source == "built"
    assert small.stats()["entries"] == 1
    assert np.array_equal(*applies(torch, e2, E, x2))
    small.get(*p1, cmaps=c1)
    assert small.last_source == "built"  # p1 was dropped
    # the dropped kernel stays alive for the operator still using it
    assert np.array_equal((a1 @ dev(torch, x1).reshape(-1)).cpu().numpy(), y_before)
    st = small.stats()
    assert (st["built"], st["memory"], st["disk"], st["entries"]) == (3, 1, 0, 1)

    # a corrupt file is rebuilt and replaced
    raw = bytearray(f1.read_bytes())
    raw[200] ^= 0x01
    f1.write_bytes(bytes(raw))
    third = msr.KernelCache(directory=str(tmp_path))
    G = third.get(*p1, cmaps=c1)
    assert third.last_source == "built"
    kf.read(str(f1))  # valid again
    assert np.array_equal(*applies(torch, G, A, x1))
    assert not [p for p in os.listdir(tmp_path) if ".tmp." in p]
