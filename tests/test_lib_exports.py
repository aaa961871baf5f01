"""The C-ABI library loads and exports every function include/cms_hip.h
declares (no device calls: runs without a GPU)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cms_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cms_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_abi():
    names = declared_functions()
    assert "cms_create" in names and "cms_generate" in names and "cms_nearest" in names


def test_library_exports_every_declared_symbol():
    from cms_cuda import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libcms_hip.so not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes prototypes cover the same set
    assert set(_lib.exported_symbols()) <= set(declared_functions())


def test_version_without_gpu():
    from cms_cuda import _lib
    lib = _lib.load()
    assert lib.cms_version() >= 100
    # cms_last_error(NULL) is always a valid string
    assert isinstance(lib.cms_last_error(None), bytes)
