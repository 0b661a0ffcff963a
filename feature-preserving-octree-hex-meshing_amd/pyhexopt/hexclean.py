"""ctypes binding of the hex-mesh cleaning passes of libhexopt.so (include/hexclean.h).

The functions carry the reference's names (grid_meshing/grid_hex_meshing.cpp:1932-2124,
global_functions.cpp:2199-2229) and return new arrays; the flags are one bool per hex, True
for a hex in the sub-mesh. Host code only: nothing here needs a GPU.
"""
import ctypes as C

import numpy as np

from . import capi

EXPORTS = ["hexclean_error", "hexclean_reorder_hex_mesh", "hexclean_tagging_uneven_element",
           "hexclean_clean_non_manifold_ve", "hexclean_drop_small_pieces", "hexclean_clean_flags",
           "hexclean_medial_surface"]

_READY = False


class HexCleanError(ValueError):
    pass


def _lib():
    global _READY
    L = capi.lib()
    if not _READY:
        vp, i64 = C.c_void_p, C.c_int64
        L.hexclean_error.restype = C.c_char_p
        L.hexclean_reorder_hex_mesh.argtypes = [i64, vp, i64, vp, vp]
        L.hexclean_reorder_hex_mesh.restype = i64
        for n in EXPORTS[2:6]:
            getattr(L, n).argtypes = [i64, i64, vp, vp]
            getattr(L, n).restype = i64
        L.hexclean_medial_surface.argtypes = [i64, i64, vp, vp, vp, vp]
        L.hexclean_medial_surface.restype = i64
        _READY = True
    return L


def _hexes(H):
    H = np.ascontiguousarray(H, dtype=np.int32)
    if H.ndim != 2 or H.shape[1] != 8:
        raise ValueError("H must be (n, 8)")
    return H


def _run(name, nV, H, flag):
    L = _lib()
    H = _hexes(H)
    f = np.ascontiguousarray(np.asarray(flag, dtype=bool), dtype=np.uint8).copy()
    if f.shape != (len(H),):
        raise ValueError("one flag per hex")
    r = getattr(L, name)(int(nV), len(H), H.ctypes.data, f.ctypes.data)
    if r < 0:
        raise HexCleanError(L.hexclean_error().decode(errors="replace"))
    return f.astype(bool)


def reorder_hex_mesh(V, H):
    """Hexes with a negative summed corner-tetrahedron volume get their corners reordered to
    3,2,1,0,7,6,5,4; returns (H', number reordered)."""
    L = _lib()
    V = np.ascontiguousarray(V, dtype=np.float64)
    H = _hexes(H).copy()
    n = C.c_int64(0)
    if L.hexclean_reorder_hex_mesh(len(V), V.ctypes.data, len(H), H.ctypes.data, C.byref(n)) < 0:
        raise HexCleanError(L.hexclean_error().decode(errors="replace"))
    return H, int(n.value)


def tagging_uneven_element(nV, H, flag):
    return _run("hexclean_tagging_uneven_element", nV, H, flag)


def clean_non_manifold_ve(nV, H, flag):
    return _run("hexclean_clean_non_manifold_ve", nV, H, flag)


def drop_small_pieces(nV, H, flag):
    return _run("hexclean_drop_small_pieces", nV, H, flag)


def clean_flags(nV, H, flag):
    """tagging_uneven_element, then (unless nothing is flagged) clean_non_manifold_ve and
    drop_small_pieces: clean_hex_mesh's flags from the inside flags signed_dis < 0."""
    return _run("hexclean_clean_flags", nV, H, flag)


def medial_surface(nV, H, flag):
    """The medial-surface tags closing clean_hex_mesh (grid_hex_meshing.cpp:1966-1981): the
    vertex flags (nV bools) and the faces on it (m x 4 vertex ids, in the reference's face
    order) of the flagged region's boundary."""
    L = _lib()
    H = _hexes(H)
    f = np.ascontiguousarray(np.asarray(flag, dtype=bool), dtype=np.uint8)
    if f.shape != (len(H),):
        raise ValueError("one flag per hex")
    v = np.zeros(int(nV), dtype=np.uint8)
    q = np.zeros((max(1, 6 * len(H)), 4), dtype=np.int32)
    n = L.hexclean_medial_surface(int(nV), len(H), H.ctypes.data, f.ctypes.data, v.ctypes.data, q.ctypes.data)
    if n < 0:
        raise HexCleanError(L.hexclean_error().decode(errors="replace"))
    return v.astype(bool), q[:n].copy()
