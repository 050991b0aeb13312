"""The C-ABI library loads without a GPU and exports exactly the entry points
include/botorch_amd.h declares, with the signatures the ctypes binding uses."""
import ctypes
import os
import re

import pytest

from botorch_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "botorch_amd.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(bta_[a-z0-9_]+)\s*\(", text))


def test_header_declares_what_binding_binds():
    assert _declared() == set(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name), name


def test_library_metadata_without_gpu():
    lib = _lib.load()
    assert lib.bta_abi_version() == _lib.ABI_VERSION == 5
    assert b"gfx950" in lib.bta_version()


def test_workspace_sizes_are_consistent():
    lib = _lib.load()
    assert lib.bta_potrf_workspace_size(256) >= 256 * 256 * 8
    assert lib.bta_mll_workspace_size(4096, 16) >= 2 * 4096 * 4096 * 8
    m = _lib.GpModel()
    m.n, m.n_pad, m.d = 2048, 2048, 12
    st = _lib.QLogNEIState()
    st.q, st.n_b, st.S, st.precision = 4, 2, 512, _lib.PREC_MIXED
    ws = lib.bta_qlognei_workspace_size(ctypes.byref(m), ctypes.byref(st), 512)
    assert ws >= 2048 * 2048 * 4 * 3
    assert lib.bta_mll_workspace_size(1, 1) == 0


def test_invalid_arguments_return_error_not_crash():
    lib = _lib.load()
    rc = lib.bta_gemm(0, 100, 64, 64, 1.0, None, 64, None, 64, 0.0, None, 0, None, 64, 1, 0, 0,
                      0, 0, None)
    assert rc == _lib.BTA_ERR_INVALID
    assert b"bta_gemm" in lib.bta_last_error()
    # the fused Cholesky + inverse refuses its operands before touching the device
    # (pointers are never dereferenced on these paths)
    jit = ctypes.c_double(0)
    ws_n = lib.bta_potrf_workspace_size(128)
    A, ws = ctypes.c_void_p(1 << 20), ctypes.c_void_p(1 << 30)
    rc = lib.bta_potrf_inverse(A, 128, 128, None, 128, 6, 1e-8, ctypes.byref(jit), ws, ws_n, None)
    assert rc == _lib.BTA_ERR_INVALID and b"null" in lib.bta_last_error()
    rc = lib.bta_potrf_inverse(A, 128, 128, ctypes.c_void_p((1 << 20) + 64 * 8), 128, 6, 1e-8,
                               ctypes.byref(jit), ws, ws_n, None)
    assert rc == _lib.BTA_ERR_INVALID and b"overlaps" in lib.bta_last_error()
    rc = lib.bta_potrf_inverse(A, 128, 128, ctypes.c_void_p((1 << 24) + 8), 128, 6, 1e-8,
                               ctypes.byref(jit), ws, ws_n, None)
    assert rc == _lib.BTA_ERR_INVALID and b"aligned" in lib.bta_last_error()
