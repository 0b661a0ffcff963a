"""ctypes binding of the mesh and feature-graph files of libhexopt.so (include/hexio.h).

The calls mirror the reference's h_io (io.h:13-31): the same files, byte for byte, with
numbers written as %g and read as sscanf reads them. Host code only: nothing here needs a GPU.
Cells are an (n, k) int array when every cell has k vertices, otherwise a (off, idx) CSR pair.
"""
import ctypes as C
from typing import NamedTuple

import numpy as np

from . import capi

TRI, QUA, HSUR, TET, HYB, HEX = 0, 1, 2, 3, 4, 5

EXPORTS = ["hexio_error", "hexio_read_vtk", "hexio_read_mesh", "hexio_read_obj", "hexio_mesh_sizes",
           "hexio_mesh_copy", "hexio_mesh_free", "hexio_write_vtk", "hexio_write_mesh", "hexio_write_obj",
           "hexio_write_matrix_obj", "hexio_read_fgraph", "hexio_fgraph_info", "hexio_fgraph_copy",
           "hexio_fgraph_free", "hexio_write_fgraph"]

_READY = False


class HexIOError(IOError):
    pass


class FgraphHeader(C.Structure):
    _fields_ = [("angle_threshold", C.c_double), ("orphan_curve", C.c_int32), ("orphan_curve_single", C.c_int32),
                ("n_corners", C.c_int32), ("n_edges", C.c_int32)]


class MeshData(NamedTuple):
    V: np.ndarray      # (nV, 3) float64
    off: np.ndarray    # (nCells + 1,) int32
    idx: np.ndarray    # int32 vertex ids, 0-based
    type: int          # TRI, QUA, TET, HYB or HEX

    @property
    def cells(self):
        """(nCells, k) array; raises when the cells have different sizes."""
        k = np.diff(self.off)
        if len(k) == 0:
            return self.idx.reshape(0, 0)
        if not np.all(k == k[0]):
            raise ValueError("cells of different sizes: use off and idx")
        return self.idx.reshape(-1, int(k[0]))


class FeatureGraph(NamedTuple):
    angle_threshold: float
    orphan_curve: bool
    orphan_curve_single: bool
    corners: np.ndarray   # (n,) int32
    pairs: np.ndarray     # (m, 2) int32


def _lib():
    global _READY
    L = capi.lib()
    if not _READY:
        vp, i32, dbl = C.c_void_p, C.c_int, C.c_double
        L.hexio_error.restype = C.c_char_p
        for n in ("hexio_read_vtk", "hexio_read_mesh", "hexio_read_obj", "hexio_read_fgraph"):
            getattr(L, n).restype = vp
            getattr(L, n).argtypes = [C.c_char_p]
        L.hexio_mesh_sizes.argtypes = [vp, vp]
        L.hexio_mesh_copy.argtypes = [vp, vp, vp, vp]
        L.hexio_mesh_free.argtypes = [vp]
        L.hexio_mesh_free.restype = None
        L.hexio_fgraph_info.argtypes = [vp, C.POINTER(FgraphHeader)]
        L.hexio_fgraph_copy.argtypes = [vp, vp, vp]
        L.hexio_fgraph_free.argtypes = [vp]
        L.hexio_fgraph_free.restype = None
        L.hexio_write_vtk.argtypes = [C.c_char_p, i32, i32, vp, i32, i32, vp, vp, vp]
        L.hexio_write_mesh.argtypes = [C.c_char_p, i32, i32, vp, i32, i32, vp, vp]
        L.hexio_write_obj.argtypes = [C.c_char_p, i32, vp, i32, i32, vp, vp]
        L.hexio_write_matrix_obj.argtypes = [C.c_char_p, i32, i32, vp, i32, i32, vp]
        L.hexio_write_fgraph.argtypes = [C.c_char_p, dbl, i32, i32, i32, vp, i32, vp]
        for n in ("hexio_mesh_sizes", "hexio_mesh_copy", "hexio_fgraph_info", "hexio_fgraph_copy",
                  "hexio_write_vtk", "hexio_write_mesh", "hexio_write_obj", "hexio_write_matrix_obj",
                  "hexio_write_fgraph"):
            getattr(L, n).restype = C.c_int
        _READY = True
    return L


def _err(L):
    return HexIOError(L.hexio_error().decode(errors="replace"))


def _path(p):
    return str(p).encode()


def _ptr(a):
    return None if a is None else a.ctypes.data


def _read_mesh(fn, path):
    L = _lib()
    h = getattr(L, fn)(_path(path))
    if not h:
        raise _err(L)
    try:
        s = (C.c_int64 * 4)()
        L.hexio_mesh_sizes(h, s)
        nV, nC, nI, t = (int(x) for x in s)
        V = np.empty((nV, 3), np.float64)
        off = np.empty(nC + 1, np.int32)
        idx = np.empty(nI, np.int32)
        L.hexio_mesh_copy(h, V.ctypes.data, off.ctypes.data, idx.ctypes.data)
    finally:
        L.hexio_mesh_free(h)
    return MeshData(V, off, idx, t)


def read_vtk(path):
    """h_io::read_hybrid_mesh_VTK (io.cpp:5)."""
    return _read_mesh("hexio_read_vtk", path)


def read_mesh(path):
    """h_io::read_hybrid_mesh_MESH (io.cpp:181)."""
    return _read_mesh("hexio_read_mesh", path)


def read_obj(path):
    """h_io::read_hybrid_mesh_OBJ (io.cpp:326), i.e. igl::readOBJ."""
    return _read_mesh("hexio_read_obj", path)


def _cells(cells):
    """(nC, k, off, idx) arrays for the C calls from an (n, k) array or an (off, idx) pair."""
    if isinstance(cells, tuple):
        off = np.ascontiguousarray(cells[0], np.int32)
        idx = np.ascontiguousarray(cells[1], np.int32)
        return len(off) - 1, 0, off, idx
    c = np.ascontiguousarray(cells, np.int32)
    if c.ndim != 2:
        raise ValueError("cells must be an (n, k) array or an (off, idx) pair")
    return c.shape[0], c.shape[1], None, c


def _V(V, dim=3):
    V = np.ascontiguousarray(V, np.float64)
    if V.ndim != 2 or (dim is not None and V.shape[1] != dim):
        raise ValueError("V must be an (n, %s) array" % dim)
    return V


def write_vtk(path, V, cells, type=HEX, boundary=None):
    """h_io::write_hybrid_mesh_VTK (io.cpp:102)."""
    L = _lib()
    V = _V(V)
    nC, k, off, idx = _cells(cells)
    b = None if boundary is None else np.ascontiguousarray(boundary, np.uint8)
    if b is not None and b.shape != (V.shape[0],):
        raise ValueError("boundary must hold one flag per vertex")
    if L.hexio_write_vtk(_path(path), type, V.shape[0], _ptr(V), nC, k, _ptr(off), _ptr(idx), _ptr(b)):
        raise _err(L)


def write_mesh(path, V, cells, type=HEX):
    """h_io::write_hybrid_mesh_MESH (io.cpp:295)."""
    L = _lib()
    V = _V(V)
    nC, k, off, idx = _cells(cells)
    if L.hexio_write_mesh(_path(path), type, V.shape[0], _ptr(V), nC, k, _ptr(off), _ptr(idx)):
        raise _err(L)


def write_obj(path, V, faces):
    """h_io::write_hybrid_mesh_OBJ (io.cpp:358)."""
    L = _lib()
    V = _V(V)
    nF, k, off, idx = _cells(faces)
    if L.hexio_write_obj(_path(path), V.shape[0], _ptr(V), nF, k, _ptr(off), _ptr(idx)):
        raise _err(L)


def write_matrix_obj(path, V, F):
    """h_io::write_mesh_OBJ (io.cpp:378) for one vertex per row of V and one face per row of F."""
    L = _lib()
    V = _V(V, dim=None)
    F = np.ascontiguousarray(F, np.int32)
    if F.ndim != 2:
        raise ValueError("F must be an (n, k) array")
    if L.hexio_write_matrix_obj(_path(path), V.shape[0], V.shape[1], _ptr(V), F.shape[0], F.shape[1], _ptr(F)):
        raise _err(L)


def read_fgraph(path):
    """h_io::read_feature_Graph_FGRAPH (io.cpp:412). Raises when the file cannot be opened
    (where the reference returns false) or is malformed."""
    L = _lib()
    g = L.hexio_read_fgraph(_path(path))
    if not g:
        raise _err(L)
    try:
        h = FgraphHeader()
        L.hexio_fgraph_info(g, C.byref(h))
        corners = np.empty(h.n_corners, np.int32)
        pairs = np.empty((h.n_edges, 2), np.int32)
        L.hexio_fgraph_copy(g, corners.ctypes.data, pairs.ctypes.data)
    finally:
        L.hexio_fgraph_free(g)
    return FeatureGraph(h.angle_threshold, bool(h.orphan_curve), bool(h.orphan_curve_single), corners, pairs)


def write_fgraph(path, corners, pairs, angle_threshold=0.0, orphan_curve=True, orphan_curve_single=True):
    """h_io::write_feature_Graph_FGRAPH (io.cpp:436): pairs are the flagged feature edges'
    vertex ids in the order to write them."""
    L = _lib()
    c = np.ascontiguousarray(corners, np.int32).reshape(-1)
    p = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
    if L.hexio_write_fgraph(_path(path), float(angle_threshold), int(bool(orphan_curve)),
                            int(bool(orphan_curve_single)), len(c), _ptr(c), len(p), _ptr(p)):
        raise _err(L)
