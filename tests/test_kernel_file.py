"""CPU checks of the kernel file format (msr_kernel_file_info needs no device):
the library reads a file made from the documented layout by an independent
numpy writer, and refuses truncated, corrupted, foreign and inconsistent
files with OSError (MSR_ERR_IO) or NotImplementedError (unknown version)."""
import os

import numpy as np
import pytest

import kernel_file as kf


@pytest.fixture()
def sample(tmp_path):
    rng = np.random.default_rng(3)
    img_shape, K, nkm = (12, 8), 3, 150
    P = K * (K + 1) // 2
    lam = (rng.normal(size=(nkm, P)) + 1j * rng.normal(size=(nkm, P))).astype(np.complex64)
    kmask = np.sort(rng.choice(4 * 12 * 8, nkm, replace=False)) + 1
    path = str(tmp_path / "k.msrk")
    kf.write(path, img_shape, lam, kmask, key=(0x0123456789ABCDEF, 0xFEDCBA9876543210))
    return path, img_shape, lam, kmask


def test_hash_properties():
    # order matters, one flipped bit changes both lanes, length is hashed
    a = bytes(range(64))
    h = kf.word_hash(a, 4)
    assert kf.word_hash(a, 4) == h
    b = bytearray(a)
    b[17] ^= 1
    h2 = kf.word_hash(bytes(b), 4)
    assert h2[0] != h[0] and h2[1] != h[1]
    assert kf.word_hash(a[8:16] + a[:8] + a[16:], 4) != h  # swapped words
    assert kf.word_hash(a + b"\0\0\0\0", 4) != h           # zero padding is not free
    assert kf.word_hash(a, 5) != h                          # seed


def test_library_reads_a_file_written_from_the_layout(msr, sample):
    path, img_shape, lam, kmask = sample
    info = msr.kernel_file_info(path, verify=True)
    assert info["img_shape"] == img_shape
    assert info["Ncoeff"] == 3 and not info["lambda_real"] and info["nkm"] == len(kmask)
    assert info["key"] == "0123456789abcdeffedcba9876543210"
    info2, lam2, km2 = kf.read(path)
    assert info2 == info
    assert np.array_equal(lam2, lam) and np.array_equal(km2, kmask)


def test_real_kernel_and_3d_header(msr, tmp_path):
    rng = np.random.default_rng(4)
    lam = rng.normal(size=(40, 10)).astype(np.float32)  # K = 4
    km = np.arange(1, 41)
    path = str(tmp_path / "r.msrk")
    kf.write(path, (4, 6, 5), lam, km)
    info = msr.kernel_file_info(path)
    assert info["img_shape"] == (4, 6, 5) and info["Ncoeff"] == 4 and info["lambda_real"]
    assert info["key"] == "0" * 32


def _damaged(src, dst, fn):
    data = bytearray(open(src, "rb").read())
    data = fn(data)
    open(dst, "wb").write(bytes(data))
    return dst


def test_damaged_files_are_refused(msr, sample, tmp_path):
    path = sample[0]
    d = str(tmp_path / "bad.msrk")

    def flip(off):
        def f(b):
            b[off] ^= 0x10
            return b
        return f

    # payload byte: the header alone still reads, the verified read fails
    _damaged(path, d, flip(128 + 8 * 150 + 33))
    assert msr.kernel_file_info(d, verify=False)["nkm"] == 150
    with pytest.raises(OSError, match="checksum"):
        msr.kernel_file_info(d, verify=True)
    # header field (nkm) and header checksum
    with pytest.raises(OSError):
        msr.kernel_file_info(_damaged(path, d, flip(57)), verify=False)
    # truncated payload, truncated header, trailing bytes, empty file
    with pytest.raises(OSError):
        msr.kernel_file_info(_damaged(path, d, lambda b: b[:-8]), verify=False)
    with pytest.raises(OSError):
        msr.kernel_file_info(_damaged(path, d, lambda b: b[:100]), verify=False)
    with pytest.raises(OSError):
        msr.kernel_file_info(_damaged(path, d, lambda b: b + b"\0" * 8), verify=False)
    with pytest.raises(OSError):
        msr.kernel_file_info(_damaged(path, d, lambda b: b[:0]), verify=False)
    # not a kernel file; missing file
    with pytest.raises(OSError):
        msr.kernel_file_info(_damaged(path, d, lambda b: b"NOTAKERN" + b[8:]), verify=False)
    with pytest.raises(OSError):
        msr.kernel_file_info(str(tmp_path / "missing.msrk"))
    assert not os.path.exists(str(tmp_path / "missing.msrk"))


def test_inconsistent_headers_are_refused(msr, tmp_path):
    rng = np.random.default_rng(5)
    lam = (rng.normal(size=(10, 6)) + 0j).astype(np.complex64)
    p = str(tmp_path / "x.msrk")
    # kmask longer than the 2N grid
    kf.write(p, (2, 2), np.zeros((17, 6), np.complex64), np.arange(1, 18))
    with pytest.raises(OSError, match="nkm"):
        msr.kernel_file_info(p)
    # P not triangular for K
    kf.write(p, (4, 4), lam[:, :5], np.arange(1, 11))
    with pytest.raises(OSError):
        msr.kernel_file_info(p)
    # a later format version (header checksum made valid)
    kf.write(p, (4, 4), lam, np.arange(1, 11))
    raw = bytearray(open(p, "rb").read())
    h = np.frombuffer(bytes(raw[:128]), dtype=kf.HEADER).copy()
    h["version"] = 2
    h["header_hash"] = kf.word_hash(h.tobytes()[:104], 3)[0]
    raw[:128] = h.tobytes()
    open(p, "wb").write(bytes(raw))
    with pytest.raises(NotImplementedError):
        msr.kernel_file_info(p)
