"""CPU checks of the C-ABI boundary: the library loads and exports every symbol
include/msr.h declares, with the binding's signature table in step with the
header; host-side argument mapping.  No compute calls (no GPU here)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT, load_package

HEADER = os.path.join(ROOT, "include", "msr.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(msr_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_expected_entry_points():
    syms = declared_symbols()
    for s in ["msr_ctx_create", "msr_normalop_create", "msr_normalop_create_from_traj",
              "msr_normalop_apply", "msr_backprojection", "msr_cg", "msr_traj_create"]:
        assert s in syms


def test_library_exports_every_declared_symbol(msr):
    lib = msr._lib.load()
    for s in declared_symbols():
        assert hasattr(lib, s), s
    assert set(msr._lib.SIGNATURES) == set(declared_symbols())


def test_library_error_path_without_device(msr):
    # ctx creation on a machine without a GPU fails cleanly with a message
    h = ctypes.c_void_p()
    code = msr._lib.load().msr_ctx_create(0, ctypes.byref(h))
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if not has_gpu:
        assert code != 0
        assert len(msr._lib.last_error()) > 0
    else:
        assert code == 0
        msr._lib.load().msr_ctx_destroy(h)


def test_error_code_mapping(msr):
    L = msr._lib
    with pytest.raises(AssertionError):
        L.check(L.MSR_ERR_ASSERTION)
    with pytest.raises(L.ArgumentError):
        L.check(L.MSR_ERR_ARGUMENT)
    with pytest.raises(L.ErrorException):
        L.check(L.MSR_ERR_ERROR)
    with pytest.raises(NotImplementedError):
        L.check(L.MSR_ERR_UNSUPPORTED)
    L.check(L.MSR_OK)


def test_host_api_rejects_host_arrays(msr):
    import numpy as np
    with pytest.raises(TypeError):
        msr.NFFTNormalOp((8, 8), np.zeros((2, 16, 2), np.float32), np.ones((1, 2), np.complex64))


def test_vector_checks_reject_mismatches_before_any_call(msr):
    """mul_, cg_, forward(out=...), the batch and the server check every vector
    the library reads or writes (device, dtype, length, contiguity) and every
    coil map's shape before a pointer crosses the boundary; exercised here on
    CPU tensors against a CPU 'device' (the checks do not touch the GPU)."""
    import torch
    cpu = torch.device("cpu")
    ok = torch.zeros(12, dtype=torch.complex64)
    assert msr._check_vec(ok, 12, "x", cpu, torch.complex64) is ok
    bad = [
        (torch.zeros(11, dtype=torch.complex64), "elements"),
        (torch.zeros(12, dtype=torch.complex128), "dtype"),
        (torch.zeros(24, dtype=torch.complex64)[::2], "contiguous"),
        (torch.zeros(12, dtype=torch.float32), "dtype"),
    ]
    for t, what in bad:
        with pytest.raises(msr.ArgumentError, match=what):
            msr._check_vec(t, 12, "x", cpu, torch.complex64)
    with pytest.raises(msr.ArgumentError, match="tensor"):
        msr._check_vec([0.0] * 12, 12, "x", cpu, torch.complex64)
    # coil maps: (Ny, Nx) row-major views of img_shape (Nx, Ny)
    msr._check_cmaps([torch.zeros(6, 4), torch.zeros(6, 4)], (4, 6))
    with pytest.raises(msr.ArgumentError, match=r"cmaps\[1\]"):
        msr._check_cmaps([torch.zeros(6, 4), torch.zeros(4, 6)], (4, 6))
