"""CPU checks of the C ABI: the library loads, exports every entry point the
header declares, and rejects invalid host arguments before touching a GPU
(MatFlat validates before any write, e.g. Factorization.Lu.cs:124-147)."""
import ctypes
import os
import re

import pytest

from matflat_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "matflat_hip.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(mfh_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "mfh_dgetrf_vbatched" in names and "mfh_dposv_vbatched" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    # and the ctypes table covers them all
    for name in declared_functions():
        if name != "mfh_last_error":
            assert name in _lib.SIGNATURES, name


def test_null_handle_rejected_as_argument_1():
    lib = _lib.load()
    assert lib.mfh_dposv_vbatched(None, 1, None, None, None, None, None, 8) == -1
    assert "handle" in _lib.last_error()
    assert lib.mfh_dgetrf_vbatched(None, 1, None, None, None, None, None, None, None, 4, 4) == -1
    assert lib.mfh_plan_execute(None) == -1
    assert lib.mfh_version() == 100
    assert lib.mfh_dgetri_vbatched(None, 1, None, None, None, None, None, 8) == -1
    assert lib.mfh_dgemv_vbatched(None, 1, None, None, None, None, None, None, None, None, None, None, 4, 4) == -1
    assert lib.mfh_set_bucket_streams(None, 2) == -1
    assert lib.mfh_plan_create_dgemm(None, 1, *([None] * 14), 4, 4, 4, None) == -1
    assert lib.mfh_plan_create_zgemm(None, 1, *([None] * 14), 4, 4, 4, None) == -1


def test_empty_batches_launch_nothing():
    """count = 0 returns OK before any pointer is looked at, as a batch of no
    matrices has nothing to check (MatFlat has no batched form; its
    per-matrix checks run per product on the device)."""
    lib = _lib.load()
    fake = ctypes.c_void_p(1)  # never dereferenced: the call returns first
    assert lib.mfh_dgetri_vbatched(fake, 0, None, None, None, None, None, 8) == 0
    assert lib.mfh_dgemv_vbatched(fake, 0, None, None, None, None, None, None, None, None, None, None, 4, 4) == 0


def test_create_without_gpu_reports_hip_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = _lib.load()
    h = ctypes.c_void_p()
    st = lib.mfh_create(ctypes.byref(h))
    assert st == _lib.MFH_ERR_HIP
    assert "HIP" in _lib.last_error()
