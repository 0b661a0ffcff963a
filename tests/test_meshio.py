"""Mesh and feature-graph files (include/hexio.h) against the reference's h_io (io.cpp).

Host code only (CPU tests). Pinned three ways:
  - tests/golden/io_files.npz holds, for seeded inputs, the bytes the reference's writers wrote
    (compiled from io.cpp) and, for those files plus hand-written ones, the arrays its readers
    made of them: the writers here must give the same bytes, the readers the same arrays;
  - where the compiled driver oracle/_ref/io_ref exists, larger seeded meshes are compared
    with it live;
  - at full mesh sizes, write -> read round trips give back the ids exactly and every
    coordinate as float("%g" % x), which is what six significant digits fix independently of
    either implementation.
"""
import os
import sys

import numpy as np
import pytest

from pyhexopt import meshio as mio

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import iocases as ic  # noqa: E402

FIX = np.load(os.path.join(HERE, "golden", "io_files.npz"))


def _names(prefix):
    return sorted({k.split("__")[1] for k in FIX.files if k.startswith(prefix + "__")})


def _case(name):
    g = lambda key: FIX["w__%s__%s" % (name, key)]  # noqa: E731
    c = dict(fmt=str(g("fmt")), type=int(g("type")), V=g("V"))
    if "w__%s__cells" % name in FIX.files:
        c["cells"] = g("cells")
    else:
        c["cells"] = (g("off"), g("idx"))
    c["boundary"] = g("boundary") if "w__%s__boundary" % name in FIX.files else None
    return c


def _write(c, path):
    fmt = c["fmt"]
    if fmt == "vtk":
        mio.write_vtk(path, c["V"], c["cells"], c["type"], c.get("boundary"))
    elif fmt == "mesh":
        mio.write_mesh(path, c["V"], c["cells"], c["type"])
    elif fmt == "obj":
        mio.write_obj(path, c["V"], c["cells"])
    elif fmt in ("mobj", "mobjc"):
        mio.write_matrix_obj(path, c["V"], c["cells"])
    else:
        raise ValueError(fmt)


def _same_V(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape
    # bit patterns, so -0.0, nan signs and denormals count
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("name", _names("w"))
def test_writer_bytes_match_reference(name, tmp_path):
    c = _case(name)
    p = str(tmp_path / name)
    _write(c, p)
    got = open(p, "rb").read()
    ref = bytes(FIX["w__%s__ref" % name])
    assert got == ref


@pytest.mark.parametrize("name", _names("g"))
def test_fgraph_writer_bytes_match_reference(name, tmp_path):
    g = lambda key: FIX["g__%s__%s" % (name, key)]  # noqa: E731
    p = str(tmp_path / name)
    mio.write_fgraph(p, g("corners"), g("pairs"), float(g("angle")), bool(g("oc")), bool(g("ocs")))
    assert open(p, "rb").read() == bytes(g("ref"))


@pytest.mark.parametrize("name", _names("r"))
def test_reader_matches_reference(name, tmp_path):
    g = lambda key: FIX["r__%s__%s" % (name, key)]  # noqa: E731
    fmt = str(g("fmt"))
    p = str(tmp_path / name)
    open(p, "wb").write(bytes(g("file")))
    if fmt == "fgraph":
        fg = mio.read_fgraph(p)
        assert int(g("ok")) == 1
        assert fg.angle_threshold == float(g("angle"))
        assert (fg.orphan_curve, fg.orphan_curve_single) == (bool(g("oc")), bool(g("ocs")))
        assert np.array_equal(fg.corners, g("corners"))
        assert np.array_equal(fg.pairs, g("pairs"))
        return
    m = {"vtk": mio.read_vtk, "mesh": mio.read_mesh, "obj": mio.read_obj}[fmt](p)
    _same_V(m.V, g("V"))
    assert np.array_equal(m.off, g("off"))
    assert np.array_equal(m.idx, g("idx"))
    if fmt == "mesh":  # the only reader of the reference that sets the type
        assert m.type == int(g("type"))


def test_fixture_covers_every_format():
    w = {str(FIX["w__%s__fmt" % n]) for n in _names("w")}
    r = {str(FIX["r__%s__fmt" % n]) for n in _names("r")}
    assert w == {"vtk", "mesh", "obj", "mobj", "mobjc"}
    assert r == {"vtk", "mesh", "obj", "fgraph"}
    assert len(_names("g")) >= 3


def test_read_types(tmp_path):
    V, H = ic.grid_hex(2)
    p = str(tmp_path / "a.vtk")
    mio.write_vtk(p, V, H)
    assert mio.read_vtk(p).type == mio.HEX
    mio.write_vtk(p, V, H[:, :4], mio.TET)
    assert mio.read_vtk(p).type == mio.TET
    q = str(tmp_path / "a.obj")
    mio.write_obj(q, V, H[:, :3])
    assert mio.read_obj(q).type == mio.TRI
    mio.write_obj(q, V, H[:, :4])
    assert mio.read_obj(q).type == mio.QUA


def _g_round(V):
    return np.array([float("%g" % x) for x in V.ravel()]).reshape(V.shape)


@pytest.mark.parametrize("fmt", ["vtk", "mesh"])
def test_round_trip_large_hex_mesh(fmt, tmp_path):
    # 48^3 = 110,592 hexes: several threads format and parse
    V, H = ic.grid_hex(48)
    r = np.random.default_rng(7)
    V = (V / 48.0 - 0.5) * 2.0 + r.normal(0, 1e-3, V.shape)
    p = str(tmp_path / ("big." + fmt))
    (mio.write_vtk if fmt == "vtk" else mio.write_mesh)(p, V, H)
    m = (mio.read_vtk if fmt == "vtk" else mio.read_mesh)(p)
    assert m.type == mio.HEX
    assert np.array_equal(m.cells, H)
    _same_V(m.V, _g_round(V))


def test_round_trip_obj_and_fgraph(tmp_path):
    r = np.random.default_rng(8)
    V = r.normal(size=(5000, 3))
    F = r.integers(0, 5000, (9000, 3)).astype(np.int32)
    p = str(tmp_path / "s.obj")
    mio.write_obj(p, V, F)
    m = mio.read_obj(p)
    assert np.array_equal(m.cells, F)
    _same_V(m.V, _g_round(V))
    q = str(tmp_path / "s.fgraph")
    corners = r.integers(0, 5000, 8).astype(np.int32)
    pairs = r.integers(0, 5000, (192, 2)).astype(np.int32)
    mio.write_fgraph(q, corners, pairs)
    fg = mio.read_fgraph(q)
    assert fg.angle_threshold == 0.0 and fg.orphan_curve and fg.orphan_curve_single
    assert np.array_equal(fg.corners, corners) and np.array_equal(fg.pairs, pairs)


def test_ragged_cells_round_trip(tmp_path):
    off, idx = ic.ragged_cells(30, [4, 5, 6, 8, 4, 8, 5], 3)
    V = np.random.default_rng(9).random((30, 3))
    p = str(tmp_path / "h.vtk")
    mio.write_vtk(p, V, (off, idx), mio.HYB)
    m = mio.read_vtk(p)
    assert m.type == mio.HYB
    assert np.array_equal(m.off, off) and np.array_equal(m.idx, idx)
    with pytest.raises(ValueError):
        m.cells


# ---- malformed input: the reference reads these silently wrong or never returns; here each
# ---- is an error with a message
BAD = {
    "vtk_no_points": ("vtk", "# vtk\n" + "x\n" * 12 + "POINTS 1 double\n0 0 0\n"),
    "vtk_short_point": ("vtk", "POINTS 2 double\n0 0 0\n1 2\nCELLS 0 0\n"),
    "vtk_no_cells": ("vtk", "POINTS 1 double\n0 0 0\n"),
    "vtk_bad_cell_size": ("vtk", "POINTS 3 double\n0 0 0\n1 0 0\n0 1 0\nCELLS 1 4\n3 0 1 2\n"),
    "vtk_id_range": ("vtk", "POINTS 4 double\n0 0 0\n1 0 0\n0 1 0\n0 0 1\nCELLS 1 5\n4 0 1 2 4\n"),
    "vtk_truncated_cells": ("vtk", "POINTS 4 double\n0 0 0\n1 0 0\n0 1 0\n0 0 1\nCELLS 2 10\n4 0 1 2 3\n"),
    "mesh_no_vertices": ("mesh", "MeshVersionFormatted 1\nDimension 3\nEnd\n"),
    "mesh_no_cells": ("mesh", "Vertices\n1\n0 0 0 0\nTriangles\n0\nEnd\n"),
    "mesh_short_hex": ("mesh", "Vertices 8\n" + "0 0 0 0\n" * 8 + "Hexahedra 1\n1 2 3 4 5 6 7\n"),
    "mesh_id_zero": ("mesh", "Vertices 4\n" + "0 0 0 0\n" * 4 + "Tetrahedra\n1\n0 1 2 3 0\n"),
    "obj_short_vertex": ("obj", "v 0 0\n"),
    "obj_mixed_vertex": ("obj", "v 0 0 0\nv 0 0 0 1\n"),
    "obj_bad_word": ("obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 x\n"),
    "obj_mixed_forms": ("obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nf 1/1 2 3\n"),
    "obj_ragged": ("obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nf 1 2 3\nf 1 2 3 4\n"),
    "obj_id_range": ("obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n"),
    "obj_leading_space": ("obj", "  v 0 0 0\n"),
    "obj_short_normal": ("obj", "vn 0 1\n"),
    "fgraph_header": ("fgraph", "0 1\n0 0\n"),
    "fgraph_truncated": ("fgraph", "0 1 1\n2 1\n5\n"),
    "fgraph_short_edge": ("fgraph", "0 1 1\n0 1\n5\n"),
}


@pytest.mark.parametrize("name", sorted(BAD))
def test_malformed_files_fail_with_message(name, tmp_path):
    fmt, text = BAD[name]
    p = str(tmp_path / name)
    open(p, "w").write(text)
    read = {"vtk": mio.read_vtk, "mesh": mio.read_mesh, "obj": mio.read_obj, "fgraph": mio.read_fgraph}[fmt]
    with pytest.raises(mio.HexIOError) as e:
        read(p)
    assert str(e.value)


def test_missing_file_and_bad_writer_input(tmp_path):
    with pytest.raises(mio.HexIOError, match="cannot open"):
        mio.read_fgraph(str(tmp_path / "none.fgraph"))
    with pytest.raises(mio.HexIOError, match="cannot open"):
        mio.read_vtk(str(tmp_path / "none.vtk"))
    V = np.zeros((4, 3))
    with pytest.raises(mio.HexIOError, match="outside"):
        mio.write_vtk(str(tmp_path / "x.vtk"), V, np.array([[0, 1, 2, 4]]), mio.TET)
    with pytest.raises(mio.HexIOError, match="cannot open"):
        mio.write_mesh(str(tmp_path / "no" / "dir.mesh"), V, np.array([[0, 1, 2, 3]]), mio.TET)


def test_empty_meshes(tmp_path):
    V = np.zeros((0, 3))
    C0 = np.zeros((0, 8), np.int32)
    p = str(tmp_path / "e.vtk")
    mio.write_vtk(p, V, C0)
    m = mio.read_vtk(p)
    assert m.V.shape == (0, 3) and len(m.idx) == 0
    q = str(tmp_path / "e.mesh")
    mio.write_mesh(q, V, C0)
    m = mio.read_mesh(q)
    assert m.V.shape == (0, 3) and len(m.idx) == 0 and m.type == mio.HEX


# ---- live comparison with the compiled reference, where it has been built ---------------------
live = pytest.mark.skipif(not ic.available(), reason="oracle/_ref/io_ref not built (reference tree absent)")


@live
@pytest.mark.parametrize("seed", [0, 1])
def test_live_writers_and_readers_against_reference(seed, tmp_path):
    r = np.random.default_rng(100 + seed)
    V, H = ic.grid_hex(12)
    V = V * 10.0 ** r.integers(-6, 7) + r.normal(0, 1, V.shape) * 10.0 ** r.integers(-8, 3, V.shape)
    bnd = (r.random(len(V)) < 0.3).astype(np.uint8)
    for fmt in ("vtk", "mesh"):
        case = dict(fmt=fmt, type=mio.HEX, V=V, cells=H, boundary=bnd if fmt == "vtk" else None)
        a, b = str(tmp_path / ("a." + fmt)), str(tmp_path / ("b." + fmt))
        ic.ref_write(case, a)
        _write(case, b)
        assert open(a, "rb").read() == open(b, "rb").read()
        ref = ic.ref_read(fmt, a)
        m = (mio.read_vtk if fmt == "vtk" else mio.read_mesh)(a)
        _same_V(m.V, ref["V"])
        assert np.array_equal(m.off, ref["off"]) and np.array_equal(m.idx, ref["idx"])
    F = r.integers(0, len(V), (500, 3)).astype(np.int32)
    case = dict(fmt="obj", type=mio.TRI, V=V, cells=F)
    a, b = str(tmp_path / "a.obj"), str(tmp_path / "b.obj")
    ic.ref_write(case, a)
    _write(case, b)
    assert open(a, "rb").read() == open(b, "rb").read()
    ref = ic.ref_read("obj", a)
    m = mio.read_obj(a)
    _same_V(m.V, ref["V"])
    assert np.array_equal(m.idx, ref["idx"])
