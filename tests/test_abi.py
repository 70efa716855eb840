"""CPU checks of the C ABI: the library loads, exports every function that
include/zgr.h declares, and fails cleanly (no crash) where no device exists."""
import ctypes
import os
import re

import pytest

import zg_rocm as Z

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "zgr.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(zgr_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_the_binding():
    names = declared_functions()
    assert "zgr_encode" in names and "zgr_amt_prove" in names
    assert set(names) == set(Z.SIGNATURES), set(names) ^ set(Z.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = Z.lib()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_error_paths_without_device():
    out = ctypes.c_void_p()
    # geometry is validated before any device call, so this never touches a GPU
    rc = Z.lib().zgr_ctx_create(0, 10, 10, 4, ctypes.byref(out))
    assert rc == Z.ERR_INVALID_ARG
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK):
        pytest.skip("a device may be present; the no-device path is not exercised")
    rc = Z.lib().zgr_ctx_create(0, 10, 10, 3, ctypes.byref(out))
    assert rc in (Z.ERR_DEVICE, Z.ERR_INVALID_ARG)
    assert Z.lib().zgr_last_error()
    # geometry validation happens before any device call
    assert Z.lib().zgr_ctx_create(0, 2, 10, 3, ctypes.byref(out)) == Z.ERR_INVALID_ARG
    assert Z.lib().zgr_ctx_create(0, 10, 10, 4, ctypes.byref(out)) == Z.ERR_INVALID_ARG
    assert Z.lib().zgr_ctx_create(0, 10, 10, 3, None) == Z.ERR_INVALID_ARG
    # null context on every entry point
    assert Z.lib().zgr_encode(None, None, 0, None) == Z.ERR_INVALID_ARG
    assert Z.lib().zgr_params_generate(None, None) == Z.ERR_INVALID_ARG


def test_geometry_constants():
    # crates/encoder/src/constants.rs:14-28
    g = Z.PRODUCTION
    assert g.max_blob_size == 32505856 and g.max_raw_data_size == 32505852
    assert g.row_encoded == 3072
    t = Z.TEST_SETTING
    assert (t.row_n, t.col_n, t.max_raw_data_size) == (64, 32, 63484)
