"""Properties every octree of OctreeGrid::subdivide(..., graded, paired) has, checked on a leaf
list alone (n x 4 rows x, y, z, extent in voxel units) with numpy:

  - the leaves tile the grid: their volumes add up to the grid's and no leaf lies inside
    another;
  - paired (octree.cpp:632-643 and :569-589): the siblings of every non-root leaf are leaves;
  - graded (octree.cpp:598-627): the parent of every non-root leaf has its 6 face and 12 edge
    neighbours at its own level, i.e. no leaf coarser than that parent covers them.

They hold for the reference's fixtures and are the full-size checks of the GPU build.
"""
import numpy as np

OFFS18 = np.array([(x, y, z) for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1)
                   if 1 <= (x != 0) + (y != 0) + (z != 0) <= 2], dtype=np.int64)
CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]],
                   dtype=np.int64)


class _Keys:
    """one int64 per (cell position, extent): positions in units of the smallest leaf extent
    (every leaf and every ancestor position is a multiple of it)"""

    def __init__(self, grid, emin):
        self.u = int(emin)
        self.G = int(np.max(grid)) // self.u + 1
        assert self.G ** 3 * 64 < 2 ** 62, "grid too fine for the key packing"

    def __call__(self, p, e):
        q = np.asarray(p, np.int64) // self.u
        lvl = np.round(np.log2(np.asarray(e, np.float64))).astype(np.int64)
        return ((q[:, 0] * self.G + q[:, 1]) * self.G + q[:, 2]) * 64 + lvl


def _is_leaf(K, keys_sorted, p, e):
    k = K(p, np.broadcast_to(np.asarray(e, np.int64), (len(p),)))
    i = np.searchsorted(keys_sorted, k)
    i = np.minimum(i, len(keys_sorted) - 1)
    return keys_sorted[i] == k


def check_octree(leaves, grid, graded=True, paired=True):
    lv = np.asarray(leaves, np.int64)
    grid = np.asarray(grid, np.int64)
    R = grid.min()
    assert len(lv) > 0
    e = lv[:, 3]
    assert np.all(e >= 1) and np.all((e & (e - 1)) == 0) and np.all(e <= R)
    assert np.all(lv[:, :3] % e[:, None] == 0)
    assert np.all(lv[:, :3] >= 0) and np.all(lv[:, :3] + e[:, None] <= grid)
    assert (e.astype(object) ** 3).sum() == int(grid[0]) * int(grid[1]) * int(grid[2]), "volumes do not tile the grid"
    K = _Keys(grid, e.min())
    keys = np.sort(K(lv[:, :3], e))
    assert len(np.unique(keys)) == len(keys), "duplicate leaves"
    # no leaf inside another: no proper ancestor of a leaf is a leaf
    a, ae = lv[:, :3].copy(), e.copy()
    while True:
        up = ae < R
        if not up.any():
            break
        a, ae = a[up], ae[up] * 2
        a = a - a % ae[:, None]
        assert not _is_leaf(K, keys, a, ae).any(), "a leaf lies inside another leaf"
    nonroot = e < R
    if paired:
        p = lv[nonroot]
        pe = p[:, 3] * 2
        base = p[:, :3] - p[:, :3] % pe[:, None]
        for c in CORNERS:
            sib = base + c * p[:, 3:4]
            assert _is_leaf(K, keys, sib, p[:, 3]).all(), "a sibling of a leaf is split (not paired)"
    if graded:
        p = lv[nonroot]
        pe = p[:, 3] * 2                    # extent of the split parent
        par = p[:, :3] - p[:, :3] % pe[:, None]
        for o in OFFS18:
            n = par + o * pe[:, None]
            inside = np.all((n >= 0) & (n < grid), axis=1)
            n, ne = n[inside], pe[inside]
            # no leaf coarser than ne may contain n
            c, ce = n.copy(), ne.copy()
            while True:
                up = ce < R
                if not up.any():
                    break
                c, ce = c[up], ce[up] * 2
                c = c - c % ce[:, None]
                assert not _is_leaf(K, keys, c, ce).any(), "a neighbour of a split cell is coarser (not graded)"
    return True


def check_mesh(out, base, voxel_size):
    """octree_mesh's export: every leaf's corners are the nodes at its corner positions, the
    nodes are exactly those positions, and V = base + voxel_size * position."""
    lv = out["leaves"].astype(np.int64)
    nodes = out["nodes"].astype(np.int64)
    H = out["H"]
    assert len(np.unique(nodes, axis=0)) == len(nodes)
    expect = lv[:, None, :3] + lv[:, None, 3:4] * CORNERS[None]
    assert np.array_equal(nodes[H], expect)
    assert len(np.unique(H)) == len(nodes)
    V = np.asarray(base, np.float64) + nodes.astype(np.float64) * voxel_size
    assert np.array_equal(out["V"].view(np.uint64), V.view(np.uint64))
    return True
