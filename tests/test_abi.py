"""The C ABI: libself_amd.so loads and exports every function include/self_amd.h declares
(no device calls)."""
import ctypes
import os
import re

import self_amd
from self_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "self_amd.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(selfamd_\w+)\s*\(", src)))


def test_every_declared_symbol_is_exported_and_bound():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared()
    assert len(names) > 40
    for n in names:
        assert hasattr(lib, n), n
        assert n in _lib.SIGNATURES, f"{n} missing from the ctypes signature table"
    assert set(_lib.SIGNATURES) == set(names)


IO_HEADER = os.path.join(os.path.dirname(HEADER), "self_amd_io.h")


def test_io_library_exports_its_header():
    if not os.path.exists(_lib.IO_LIB_PATH):
        import pytest
        pytest.skip("libself_amd_io.so not built (no HDF5 at build time)")
    src = re.sub(r"/\*.*?\*/", "", open(IO_HEADER).read(), flags=re.S)
    names = sorted(set(re.findall(r"\b(selfamd_\w+)\s*\(", src)))
    lib = ctypes.CDLL(_lib.IO_LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert set(_lib.IO_SIGNATURES) == set(names)


def test_version_and_error_channel():
    lib = self_amd.load()
    assert lib.selfamd_version() == 1
    rc = lib.selfamd_mesh2d_structured(0, 0, 0.0, 1.0, 0.0, 1.0, 0, 0, 0, None, None, None)
    assert rc == 1
    assert b"nx=0" in lib.selfamd_last_error()
