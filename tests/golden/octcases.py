"""Wrapper of the compiled-reference octree driver (oracle/_ref/octree_ref, built from the
reference's grid_meshing/octree.cpp) and the surfaces the octree parity tests subdivide
around. Used in the development container to write tests/golden/octree_*.npz and, where the
driver exists, by the tests directly."""
import os
import struct
import subprocess
import tempfile

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OCT_BIN = os.path.join(REPO, "oracle", "_ref", "octree_ref")


def available():
    return os.path.exists(OCT_BIN)


def ref_octree(V, F, grid, origin, voxel_size, mesh_transform, stop_extent, graded=True, paired=True,
               refine=None):
    """dict(leaves (n, 4) x y z extent in cell order, nodes (m, 3) positions in node order,
    corners (n, 8) node ids, graded, paired: the reference's own is2to1Graded / isPaired).
    refine = (stop2, cells (k, 4)): then the targeted subdivide of those leaves at stop2."""
    V = np.ascontiguousarray(V, np.float64)
    F = np.ascontiguousarray(F, np.int32)
    with tempfile.TemporaryDirectory() as td:
        pin, pout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(pin, "wb") as f:
            f.write(struct.pack("<8i", *grid, stop_extent, int(graded), int(paired), len(V), len(F)))
            f.write(struct.pack("<7d", *origin, voxel_size, *mesh_transform))
            f.write(V.tobytes())
            f.write(F.tobytes())
            if refine is not None:
                cells = np.ascontiguousarray(refine[1], np.int32).reshape(-1, 4)
                f.write(struct.pack("<2i", int(refine[0]), len(cells)))
                f.write(cells.tobytes())
        subprocess.run([OCT_BIN, pin, pout], check=True, timeout=600, stdout=subprocess.DEVNULL)
        raw = open(pout, "rb").read()
    nl, nn, g, p = struct.unpack_from("<4i", raw, 0)
    a = np.frombuffer(raw, "<i4", offset=16)
    leaves = a[:4 * nl].reshape(nl, 4).copy()
    nodes = a[4 * nl:4 * nl + 3 * nn].reshape(nn, 3).copy()
    corners = a[4 * nl + 3 * nn:].reshape(nl, 8).copy()
    return dict(leaves=leaves, nodes=nodes, corners=corners, graded=g, paired=p)


def mesher_frame(V, num_voxels):
    """origin, voxel_size, grid size and mesh_transform as octree_mesh sets them up
    (grid_meshing/grid_hex_meshing.cpp:463-497, padding 0, scaffold type 1)."""
    lo, hi = V.min(0), V.max(0)
    center = (lo + hi) / 2
    extent = hi - lo
    voxel_size = max(extent) / num_voxels
    origin = lo.copy()
    grid = [next_pow2(int(np.ceil(e / voxel_size))) for e in extent]
    origin_max = origin + voxel_size * np.array(grid, np.float64)
    origin_center = (origin_max + origin) * 0.5
    mesh_transform = center - origin_center
    return origin, voxel_size, grid, mesh_transform


def next_pow2(x):
    """grid_meshing/voxelization.h:394-402 (the smallest power of two >= x, for x >= 1)."""
    x -= 1
    for s in (1, 2, 4, 8, 16):
        x |= x >> s
    return x + 1
