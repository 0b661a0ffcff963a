"""clean_hex_mesh's flag passes (include/hexclean.h, host code) against the Python oracle
(oracle/hexclean_oracle.py) and against meshes whose cleaned flags follow from the
reference's definitions (grid_meshing/grid_hex_meshing.cpp:1932-2124). CPU only.

Structured grids number vertices and hexes x fastest, corners in the reference's order
(0..3 at z, 4..7 at z+1, counter-clockwise seen from +z), so every hex is positively oriented.
"""
import os
import sys

import numpy as np
import pytest

from pyhexopt import hexclean

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import hexclean_oracle as O  # noqa: E402

CORN = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])


def grid(nx, ny, nz):
    """(V, H) of an nx x ny x nz cell grid; hex id = x + nx * (y + ny * z)."""
    vx, vy = nx + 1, ny + 1
    z, y, x = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    V = np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.float64)
    cz, cy, cx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    c = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], 1)
    p = c[:, None, :] + CORN[None]
    H = (p[..., 0] + vx * (p[..., 1] + vy * p[..., 2])).astype(np.int32)
    return V, H


def cells(shape, on):
    nx, ny, nz = shape
    f = np.zeros(nx * ny * nz, bool)
    for x, y, z in on:
        f[x + nx * (y + ny * z)] = True
    return f


def block(x0, x1, y0, y1, z0, z1):
    return [(x, y, z) for z in range(z0, z1) for y in range(y0, y1) for x in range(x0, x1)]


def both(name, V, H, flag):
    got = getattr(hexclean, name)(len(V), H, flag)
    want = np.array(getattr(O, name)(len(V), H.tolist(), flag.tolist()), bool)
    assert np.array_equal(got, want)
    return got


# ---- meshes whose result follows from the definitions -------------------------------------

def test_spike_on_a_block_is_removed():
    # the spike has its top on the grid boundary, 4 sides against unflagged hexes and its
    # bottom against the block: 5 boundary faces and 1 inner face
    V, H = grid(5, 5, 5)
    blk = cells((5, 5, 5), block(1, 4, 1, 4, 1, 4))
    f = blk | cells((5, 5, 5), [(2, 2, 4)])
    assert np.array_equal(both("tagging_uneven_element", V, H, f), blk)
    assert np.array_equal(both("clean_flags", V, H, f), blk)


def test_pit_bottom_is_filled_in_hex_order():
    # the pit's lower hex (lower id) sees 5 flagged or boundary faces and its unflagged upper
    # neighbour: it is flagged; the upper hex, visited later, then has 6 and stays out
    V, H = grid(3, 3, 4)
    f = ~cells((3, 3, 4), [(1, 1, 2), (1, 1, 3)])
    want = ~cells((3, 3, 4), [(1, 1, 3)])
    assert np.array_equal(both("tagging_uneven_element", V, H, f), want)


def test_hexes_sharing_only_an_edge_lose_the_first_boundary_face_hex():
    # 2x2x1 grid, hexes 0 and 3 meet along the edge of vertices 4 and 13, which has 4
    # boundary faces; the lowest face id is the smallest sorted vertex quadruple,
    # hex 0's +x face (1, 4, 10, 13), so hex 0 goes
    V, H = grid(2, 2, 1)
    f = cells((2, 2, 1), [(0, 0, 0), (1, 1, 0)])
    assert np.array_equal(both("tagging_uneven_element", V, H, f), f)
    assert np.array_equal(both("clean_non_manifold_ve", V, H, f), [False, False, False, True])
    assert np.array_equal(both("clean_flags", V, H, f), [False, False, False, True])


def test_hexes_sharing_only_a_vertex_keep_the_first_faces_fan():
    # 2x2x2 grid, hexes 0 and 7 meet at vertex 13: two fans of 3 boundary faces; the first
    # boundary face there is hex 0's (1, 4, 10, 13), so hex 7's fan is cut
    V, H = grid(2, 2, 2)
    f = cells((2, 2, 2), [(0, 0, 0), (1, 1, 1)])
    want = cells((2, 2, 2), [(0, 0, 0)])
    assert np.array_equal(both("clean_non_manifold_ve", V, H, f), want)
    assert np.array_equal(both("clean_flags", V, H, f), want)


def test_smaller_piece_is_dropped():
    V, H = grid(5, 2, 2)
    small = cells((5, 2, 2), [(0, 0, 0)])
    big = cells((5, 2, 2), block(3, 5, 0, 2, 0, 2))
    assert np.array_equal(both("drop_small_pieces", V, H, small | big), big)
    assert np.array_equal(both("clean_flags", V, H, small | big), big)


def test_equal_pieces_keep_the_first():
    V, H = grid(5, 1, 1)
    f = cells((5, 1, 1), [(0, 0, 0), (4, 0, 0)])
    assert np.array_equal(both("drop_small_pieces", V, H, f), cells((5, 1, 1), [(0, 0, 0)]))


def test_nothing_inside_stays_empty():
    V, H = grid(3, 3, 3)
    f = np.zeros(len(H), bool)
    assert not both("clean_flags", V, H, f).any()
    assert not both("drop_small_pieces", V, H, f).any()


def test_manifold_block_is_untouched():
    V, H = grid(4, 4, 4)
    f = cells((4, 4, 4), block(1, 3, 1, 3, 0, 4))
    for name in ("tagging_uneven_element", "clean_non_manifold_ve", "drop_small_pieces", "clean_flags"):
        assert np.array_equal(both(name, V, H, f), f)


def test_reorder_flips_only_inverted_hexes():
    V, H = grid(2, 1, 1)
    H2 = H.copy()
    H2[1] = H[1][[4, 5, 6, 7, 0, 1, 2, 3]]          # mirrored through its mid-plane
    out, n = hexclean.reorder_hex_mesh(V, H2)
    assert n == 1
    assert np.array_equal(out[0], H[0])
    assert np.array_equal(out[1], H2[1][[3, 2, 1, 0, 7, 6, 5, 4]])
    ref, m = O.reorder_hex_mesh(V.tolist(), H2.tolist())
    assert m == 1 and np.array_equal(out, ref)
    vol = sum(O._det(*(V[out[1][k]] for k in t)) for t in O.HEX_TET)
    assert vol > 0


def test_errors_are_reported():
    V, H = grid(1, 1, 1)
    bad = H.copy()
    bad[0, 0] = 99
    with pytest.raises(hexclean.HexCleanError, match="out of range"):
        hexclean.clean_flags(len(V), bad, [True])
    bad = H.copy()
    bad[0, 1] = bad[0, 0]
    with pytest.raises(hexclean.HexCleanError, match="repeated"):
        hexclean.tagging_uneven_element(len(V), bad, [True])
    dup = np.concatenate([H, H, H])
    with pytest.raises(hexclean.HexCleanError, match="more than two"):
        hexclean.drop_small_pieces(len(V), dup, [True, True, True])


# ---- random blobs: library against the oracle, in shuffled numberings -----------------------

def manifold_and_connected(nv, H, flag):
    """the cleaned sub-mesh: every boundary edge has 2 boundary faces, the boundary faces
    at every vertex form one edge-connected fan, and the hexes form one face-connected piece"""
    m, _ = O.reindex(H, flag)
    for e in range(len(m.Ebnd)):
        if m.Ebnd[e]:
            assert sum(m.Fbnd[f] for f in m.Efs[e]) == 2
    for v in range(m.nv):
        fs = {f for f in m.Vfs[v] if m.Fbnd[f]}
        if not fs:
            continue
        seen, todo = set(), [min(fs)]
        while todo:
            f = todo.pop()
            if f in seen:
                continue
            seen.add(f)
            todo += [g for e in m.Fes[f] for g in m.Efs[e] if g in fs]
        assert seen == fs
    seen, todo = set(), [0]
    while todo:
        h = todo.pop()
        if h in seen:
            continue
        seen.add(h)
        todo += [g for f in m.Hfs[h] for g in m.Fhs[f]]
    assert len(seen) == len(m.Hvs)


@pytest.mark.parametrize("seed", range(8))
def test_random_blobs_match_the_oracle(seed):
    rng = np.random.default_rng(seed)
    n = 6 if seed < 4 else 8 if seed < 6 else 12
    V, H = grid(n, n, n)
    # smoothed noise gives blobs with spikes, pits, thin bridges and separate pieces
    occ = rng.random((n, n, n)) < 0.5
    sm = occ.astype(int)
    for ax in range(3):
        sm = sm + np.roll(occ, 1, ax) + np.roll(occ, -1, ax)
    flag = (sm >= 4).transpose(2, 1, 0).ravel()      # index x fastest
    # shuffle vertex ids and hex order: the passes must follow the numbering, not the grid
    if seed % 2:
        pv = rng.permutation(len(V))
        V = V[np.argsort(pv)]
        H = pv[H].astype(np.int32)
        ph = rng.permutation(len(H))
        H, flag = H[ph], flag[ph]
    for name in ("tagging_uneven_element", "clean_non_manifold_ve", "drop_small_pieces"):
        both(name, V, H, flag)
    out = both("clean_flags", V, H, flag)
    if out.any():
        manifold_and_connected(len(V), H.tolist(), out.tolist())


def test_planted_oracle_mistake_is_caught():
    # keeping the last largest piece instead of the first changes the equal-pieces result
    V, H = grid(5, 1, 1)
    f = cells((5, 1, 1), [(0, 0, 0), (4, 0, 0)])
    got = hexclean.drop_small_pieces(len(V), H, f)
    m, hmap = O.reindex(H.tolist(), f.tolist())
    last = [False] * len(f)
    last[hmap[-1]] = True
    assert not np.array_equal(got, last)
