"""Cases for the mesh and feature-graph file parity tests, and the wrapper of the
compiled-reference file driver (oracle/_ref/io_ref, built from the reference's io.cpp).

The writer cases are seeded arrays; the reader cases are the files the reference writes for
them plus hand-written files in the forms the reference's readers accept (CRLF line ends,
tabs, two-line section headers, OBJ texture/normal forms, negative OBJ ids, hexadecimal and
octal .fgraph integers). make_golden.py stores the reference's output for every case in
tests/golden/io_files.npz; the tests compare against that file, and against the driver itself
where it has been built.
"""
import os
import struct
import subprocess
import tempfile

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
IO_BIN = os.path.join(REPO, "oracle", "_ref", "io_ref")

TRI, QUA, HSUR, TET, HYB, HEX = 0, 1, 2, 3, 4, 5


def available():
    return os.path.exists(IO_BIN)


# ---- seeded inputs ---------------------------------------------------------------------------
SPECIALS = [0.0, -0.0, 1.0, -1.0, 0.1, 1e-5, 123456.5, 1234567.0, 0.1234565, 9.999995, 99999.95, 1e300, -1e-300,
            5e-324, 2.5, 3.5, 1e6, 1e-4, 0.0001234565, float("inf"), float("-inf"), float("nan"),
            -float("nan"), 4294967296.0]


def tricky_points(n, seed, dim=3):
    """Coordinates over many decades plus values at %g's rounding and format edges."""
    r = np.random.default_rng(seed)
    V = r.uniform(-1, 1, (n, dim)) * 10.0 ** r.integers(-7, 9, (n, dim))
    flat = V.reshape(-1)
    k = min(len(SPECIALS), flat.size)
    flat[:k] = SPECIALS[:k]
    return V


def grid_hex(n):
    g = np.arange(n + 1)
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    V = np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.float64)
    vid = lambda i, j, k: (i * (n + 1) + j) * (n + 1) + k  # noqa: E731
    H = []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                H.append([vid(i, j, k), vid(i + 1, j, k), vid(i + 1, j + 1, k), vid(i, j + 1, k),
                          vid(i, j, k + 1), vid(i + 1, j, k + 1), vid(i + 1, j + 1, k + 1), vid(i, j + 1, k + 1)])
    return V, np.array(H, np.int32)


def random_cells(nV, nC, k, seed):
    r = np.random.default_rng(seed)
    return np.stack([r.choice(nV, k, replace=False) for _ in range(nC)]).astype(np.int32)


def ragged_cells(nV, sizes, seed):
    r = np.random.default_rng(seed)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    idx = np.concatenate([r.choice(nV, s, replace=False) for s in sizes]).astype(np.int32)
    return off, idx


def write_cases():
    """name -> dict(fmt, type, V, cells or (off, idx), boundary)"""
    cases = {}
    V, H = grid_hex(3)
    V = V + tricky_points(len(V), 1) * 1e-3
    V[:8] = tricky_points(8, 2)
    bnd = (np.random.default_rng(3).random(len(V)) < 0.4).astype(np.uint8)
    cases["hex_vtk"] = dict(fmt="vtk", type=HEX, V=V, cells=H, boundary=bnd)
    cases["hex_mesh"] = dict(fmt="mesh", type=HEX, V=V, cells=H)
    Vt = tricky_points(20, 4)
    cases["tet_vtk"] = dict(fmt="vtk", type=TET, V=Vt, cells=random_cells(20, 30, 4, 5))
    cases["tet_mesh"] = dict(fmt="mesh", type=TET, V=Vt, cells=random_cells(20, 30, 4, 5))
    Vs = tricky_points(25, 6)
    cases["tri_vtk"] = dict(fmt="vtk", type=TRI, V=Vs, cells=random_cells(25, 40, 3, 7))
    cases["qua_vtk"] = dict(fmt="vtk", type=QUA, V=Vs, cells=random_cells(25, 20, 4, 8))
    cases["hsur_vtk"] = dict(fmt="vtk", type=HSUR, V=Vs, cells=random_cells(25, 20, 4, 9))
    cases["hyb_vtk"] = dict(fmt="vtk", type=HYB, V=Vs, cells=ragged_cells(25, [3, 7, 4, 5, 6, 3, 8], 10))
    cases["hyb4568_vtk"] = dict(fmt="vtk", type=HYB, V=Vs, cells=ragged_cells(25, [4, 5, 6, 8, 8, 4], 11))
    cases["tri_mesh"] = dict(fmt="mesh", type=TRI, V=Vs, cells=random_cells(25, 40, 3, 7))
    cases["hsur_mesh"] = dict(fmt="mesh", type=HSUR, V=Vs, cells=random_cells(25, 20, 4, 12))
    cases["qua_mesh"] = dict(fmt="mesh", type=QUA, V=Vs, cells=random_cells(25, 20, 4, 12))
    cases["tri_obj"] = dict(fmt="obj", type=TRI, V=Vs, cells=random_cells(25, 40, 3, 13))
    cases["qua_obj"] = dict(fmt="obj", type=QUA, V=Vs, cells=random_cells(25, 20, 4, 14))
    cases["hyb_obj"] = dict(fmt="obj", type=HYB, V=Vs, cells=ragged_cells(25, [3, 4, 5, 9], 15))
    cases["tri_mobj"] = dict(fmt="mobj", type=TRI, V=Vs, cells=random_cells(25, 40, 3, 16))
    cases["tri_mobjc"] = dict(fmt="mobjc", type=TRI, V=Vs, cells=random_cells(25, 40, 3, 16))
    cases["uv_mobj"] = dict(fmt="mobj", type=TRI, V=tricky_points(12, 17, dim=2), cells=random_cells(12, 9, 3, 18))
    return cases


def fgraph_cases():
    r = np.random.default_rng(20)
    return {
        "cube_fgraph": dict(angle=0.0, oc=1, ocs=1, corners=np.arange(8, dtype=np.int32) * 17,
                            pairs=r.integers(0, 1538, (24, 2)).astype(np.int32),
                            other=r.integers(0, 1538, (30, 2)).astype(np.int32)),
        "odd_fgraph": dict(angle=30.123456789, oc=0, ocs=1, corners=np.array([5, -1, 2147483647], np.int32),
                           pairs=np.array([[0, 1], [-1, 7], [2147483647, 3]], np.int32),
                           other=np.zeros((0, 2), np.int32)),
        "empty_fgraph": dict(angle=1e-7, oc=1, ocs=0, corners=np.zeros(0, np.int32),
                             pairs=np.zeros((0, 2), np.int32), other=np.array([[1, 2]], np.int32)),
    }


# ---- hand-written files the reference's readers accept -----------------------------------------
def handwritten_files():
    """name -> (format, bytes)"""
    f = {}
    f["variant.vtk"] = ("vtk", (
        "# vtk DataFile Version 3.0\r\n"
        "written by hand\r\n"
        "ASCII\r\n"
        "\r\n"
        "DATASET UNSTRUCTURED_GRID\r\n"
        "POINTS 9 float\r\n"
        "0 0 0\r\n"
        "1.5\t0 0\r\n"
        "  1 1 0  \r\n"
        "0 1 0 7\r\n"
        "0 0 1\r\n"
        "1 0 1\r\n"
        "1 1 1\r\n"
        "0 1 1\r\n"
        "-2.5e-3 +4 0x1p-2\r\n"
        "\r\n"
        "METADATA\r\n"
        "CELLS 4 29\r\n"
        "8 0 1 2 3 4 5 6 7\r\n"
        "4 0 1 2 8 99\r\n"
        "5\t0 1 2 3 8\r\n"
        " 6 0 1 2 3 4 5 \r\n"
        "CELL_TYPES 4\r\n12\r\n10\r\n14\r\n13\r\n").encode())
    f["variant1.mesh"] = ("mesh", (
        "MeshVersionFormatted 2\n"
        "Dimension\n3\n"
        "Vertices\n"
        "9\n"
        "0 0 0 1\n1 0 0 1\n1 1 0 1\n0 1 0 1\n0 0 1 2\n1 0 1 2\n1 1 1 2\n0 1 1 2\n"
        "0.5 0.5 -1.25e1 3\n"
        "Edges\n1\n1 2 0\n"
        "Triangles 1\n1 2 3 0\n"
        "Hexahedra 2\n"
        "1 2 3 4 5 6 7 8 0\n"
        "\t1  2 3 4 5 6 7 9 4\n"
        "End\n").encode())
    f["variant2.mesh"] = ("mesh", (
        "MeshVersionFormatted 1\r\n"
        "Dimension 3\r\n"
        "Vertices 5\r\n"
        "0 0 0 0\r\n1 0 0 0\r\n0 1 0 0\r\n0 0 1 0\r\n1 1 1 0\r\n"
        "Tetrahedra\r\n"
        "2\r\n"
        "1 2 3 4 1\r\n"
        "2 3 4 5 1\r\n"
        "End").encode())
    f["variant.obj"] = ("obj", (
        "# a comment\n"
        "mtllib things.mtl\n"
        "o thing\n"
        "v 0 0 0\n"
        "v 1 0 0 \n"
        "v 0 1 0\n"
        "v 0 0 1.25e-1\n"
        "vt 0 0\n"
        "vt 1 0\n"
        "vt 0 1 0\n"
        "vn 0 0 1\n"
        "g group\n"
        "usemtl material\n"
        "s off\n"
        "f 1/1/1 2/2/1 3/3/1\n"
        "f -4//1 -2//1 -1//1\n"
        "f 1/1 3/3 4/2\n"
        "f 2 3 4\n"
        "\n"
        "l 1 2\n").encode())
    f["w4.obj"] = ("obj", (
        "v 0 0 0 1\r\nv 1 0 0 1\r\nv 0 1 0 0.5\r\nv 1 1 0 1\r\n"
        "f 1 2 4 3\r\nf 4 3 2 1\r\n").encode())
    f["variant.fgraph"] = ("fgraph", (
        "30.5 0 1\r\n"
        "3 3\r\n"
        "0x10\r\n"
        "010\r\n"
        "4294967295\r\n"
        "1 2\r\n"
        " 0x1f 7\r\n"
        "-1 3 9\r\n").encode())
    return f


# ---- the compiled-reference driver -------------------------------------------------------------
def _cells_arrays(cells):
    if isinstance(cells, tuple):
        off, idx = (np.asarray(a, np.int32) for a in cells)
        return len(off) - 1, 1, 0, off, idx
    c = np.asarray(cells, np.int32)
    return c.shape[0], 0, c.shape[1], None, c.reshape(-1)


def ref_write(case, out_path):
    """Writes case's file with the reference's writer."""
    with tempfile.TemporaryDirectory() as td:
        pin = os.path.join(td, "in.bin")
        with open(pin, "wb") as f:
            if "fmt" not in case:  # feature graph
                f.write(struct.pack("<diiiii", case["angle"], case["oc"], case["ocs"], len(case["corners"]),
                                    len(case["pairs"]), len(case["other"])))
                for a in (case["corners"], case["pairs"], case["other"]):
                    f.write(np.ascontiguousarray(a, "<i4").tobytes())
                fmt = "fgraph"
            else:
                V = np.asarray(case["V"], np.float64)
                nC, ragged, k, off, idx = _cells_arrays(case["cells"])
                b = case.get("boundary")
                f.write(struct.pack("<8i", case["type"], V.shape[0], V.shape[1], nC, ragged, k, len(idx),
                                    int(b is not None)))
                f.write(np.ascontiguousarray(V, "<f8").tobytes())
                if ragged:
                    f.write(off.astype("<i4").tobytes())
                f.write(idx.astype("<i4").tobytes())
                if b is not None:
                    f.write(np.asarray(b, np.uint8).tobytes())
                fmt = case["fmt"]
        subprocess.run([IO_BIN, "write", fmt, pin, out_path], check=True)


def ref_read(fmt, path):
    """The reference reader's result: dict(V, off, idx, type) or, for .fgraph, dict(ok, angle,
    oc, ocs, corners, pairs)."""
    with tempfile.TemporaryDirectory() as td:
        pout = os.path.join(td, "out.bin")
        subprocess.run([IO_BIN, "read", fmt, path, pout], check=True, timeout=60)
        raw = open(pout, "rb").read()
    if fmt == "fgraph":
        ok, = struct.unpack_from("<i", raw, 0)
        angle, = struct.unpack_from("<d", raw, 4)
        oc, ocs, nc, ne = struct.unpack_from("<4i", raw, 12)
        a = np.frombuffer(raw, "<i4", offset=28)
        return dict(ok=ok, angle=angle, oc=oc, ocs=ocs, corners=a[:nc].copy(), pairs=a[nc:].reshape(ne, 2).copy())
    t, nV, nC, nI = struct.unpack_from("<4i", raw, 0)
    V = np.frombuffer(raw, "<f8", count=3 * nV, offset=16).reshape(nV, 3).copy()
    rest = np.frombuffer(raw, "<i4", offset=16 + 24 * nV)
    return dict(V=V, off=rest[:nC + 1].copy(), idx=rest[nC + 1:nC + 1 + nI].copy(), type=t)
