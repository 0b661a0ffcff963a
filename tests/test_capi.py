"""The C ABI library loads and exports every function include/*.h declares (CPU only:
no compute call is made)."""
import ctypes
import os
import re

from pyhexopt import capi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header, prefix="hexopt"):
    src = open(os.path.join(REPO, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(%s_[a-z_0-9]+)\s*\(" % prefix, src)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    names = declared("hexopt.h")
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(names) == sorted(capi.EXPORTS)


def test_library_exports_every_file_format_symbol():
    from pyhexopt import meshio
    lib = ctypes.CDLL(capi.LIB_PATH)
    names = declared("hexio.h", "hexio")
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(names) == sorted(meshio.EXPORTS)


def test_library_exports_every_cleaning_symbol():
    from pyhexopt import hexclean
    lib = ctypes.CDLL(capi.LIB_PATH)
    names = declared("hexclean.h", "hexclean")
    assert len(names) == 7
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(names) == sorted(hexclean.EXPORTS)


def test_every_header_is_checked():
    assert sorted(os.listdir(os.path.join(REPO, "include"))) == ["hexclean.h", "hexio.h", "hexopt.h"]


def test_create_fails_cleanly_on_bad_mesh():
    lib = capi.lib()
    import numpy as np
    V = np.zeros((4, 3))
    H = np.full((1, 8), 7, dtype=np.int32)   # corner index out of range
    h = lib.hexopt_create(4, V.ctypes.data, 1, H.ctypes.data, 0)
    assert not h
    assert lib.hexopt_create_error().decode() != ""
