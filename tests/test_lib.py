"""The C-ABI library loads and exports every symbol include/hfpx3d.h
declares (CPU: no device call is made)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "hfpx3d.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(hfpx_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_declared_symbols():
    from hfpx3d_amd import _lib
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n


def test_binding_signatures_cover_header():
    from hfpx3d_amd import _lib
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_create_without_device_fails_cleanly():
    import ctypes
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from hfpx3d_amd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.hfpx_create(ctypes.byref(h), 0, None) == _lib.HFPX_ERR_HIP
    assert lib.hfpx_status_string(_lib.HFPX_ERR_SINGULAR) == b"matrix is singular"


def test_oracle_library_builds_and_loads():
    import oracle_lib
    o = oracle_lib.oracle()
    assert o is not None
