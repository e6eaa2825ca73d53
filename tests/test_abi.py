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
