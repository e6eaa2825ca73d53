"""The C-ABI library loads without a GPU and exports exactly what include/cmtip_amd.h
declares; the ctypes table covers every declaration."""
import os
import re
import subprocess

from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "cmtip_amd.h")
LIB = os.path.join(ROOT, "cmtip_amd", "libcmtip_amd.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(cmtip_[A-Za-z0-9_]+)\s*\(", text))


def test_header_symbols_exported():
    if not os.path.exists(LIB):
        import cmtip_amd.build as b
        b.build()
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (cmtip_[A-Za-z0-9_]+)", out))
    missing = declared() - exported
    assert not missing, f"declared but not exported: {sorted(missing)}"


def test_ctypes_table_matches_header():
    from cmtip_amd import _lib
    assert set(_lib.SIGNATURES) == declared()
    L = _lib.load()
    for name in _lib.SIGNATURES:
        assert getattr(L, name) is not None
    assert L.cmtip_version() >= 100
    assert L.cmtip_last_error() == b""


def test_product_refuses_without_gpu():
    import torch
    if torch.cuda.is_available():
        return
    from cmtip_amd import _lib
    import pytest
    with pytest.raises(RuntimeError):
        _lib.lib()


def test_argument_errors_need_no_gpu():
    """Bad arguments are rejected with CMTIP_ERR_ARG before any device work, and the
    message is kept for cmtip_last_error; valid-looking null-gamma calls pass the checks
    (they stop at the first HIP call only where a GPU is missing)."""
    import ctypes
    from cmtip_amd import _lib
    L = _lib.load()
    ERR_ARG = 1
    h = ctypes.c_void_p()
    assert L.cmtip_nufft_makeplan(3, 10, -1, 1e-6, ctypes.byref(h)) == ERR_ARG
    assert b"bad arguments" in L.cmtip_last_error()
    assert L.cmtip_nufft_makeplan(2, 0, -1, 1e-6, ctypes.byref(h)) == ERR_ARG
    assert L.cmtip_nufft_makeplan(2, 10, 0, 1e-6, ctypes.byref(h)) == ERR_ARG
    assert L.cmtip_nufft_setpts(None, 1, None, None, None, None) == ERR_ARG
    assert L.cmtip_nufft_setpts_rotated(None, None, 1, None, 1, 1.0, None, None) == ERR_ARG  # null plan
    assert L.cmtip_acsolver_create(10, 15, ctypes.byref(h)) == ERR_ARG  # odd L
    assert L.cmtip_acsolver_create(10, 18, ctypes.byref(h)) == ERR_ARG  # L < 2M
    assert L.cmtip_match_topk_bf16(None, 10, 0, None, 10, 16, None, None, None, 17, None, None) == ERR_ARG  # k > 16
    assert b"k" in L.cmtip_last_error()
    assert L.cmtip_match_argmin_f32(None, 10, 0, None, 10, 16, None, None, None, None, None) == ERR_ARG
    assert L.cmtip_match_argmin_f32(None, 10, -1, None, 10, 16, None, None, None, None, None) == ERR_ARG
    # zero-sized problems are no-ops that succeed without touching the device
    assert L.cmtip_match_argmin_f32(None, 0, 0, None, 0, 16, None, None, None, None, None) == 0
    assert L.cmtip_rotate_coords(None, 0, None, 0, 1.0, None, None, None, None) == 0
    assert L.cmtip_center_rows(None, 0, 16, None, None, None, None, None, None, None) == 0
