"""GPU occupancy (hexopt_occ_* and hexopt_hex_inside) against the oracle, bit for bit: the
subdivision predicate on random cells and through a level-by-level subdivision, voxel and
octree-cell ray parity on closed, doubled, inward, nested and open surfaces (including
columns through vertices and edges, and columns with more crossings than the kernel caches),
and the inside test of hexes."""
import numpy as np
import pytest

import oracle_py as O
from occ_cases import cube_aligned_grid, merge, random_cells, sphere
from pyhexopt import meshgen as mg

pytestmark = pytest.mark.gpu


def _torus():
    V, F = mg.torus_surface(R=1.0, r=0.35, nu=60, nv=30)
    return V, F.astype(np.int32)


@pytest.mark.parametrize("stop_extent", [1, 2])
def test_classify_cells_vs_oracle(stop_extent):
    from pyhexopt import capi
    V, F = _torus()
    s = capi.Surface(V, F)
    rng = np.random.default_rng(7)
    cells = random_cells(rng, 20000, 64, extents=(1, 2, 4, 8, 16, 32))
    o, vs, mt = np.array([-1.5, -1.5, -0.8]), 3.0 / 64, np.array([0.003, -0.01, 0.02])
    got = s.classify_cells(cells, o, vs, mt, stop_extent)
    ref = O.cell_facet_overlap(cells, V, F, o, vs, mt, stop_extent)
    assert (got == ref).all() and 0 < ref.sum() < ref.size
    # without mesh_transform (compute_octree's predicate)
    got = s.classify_cells(cells, o, vs, None, stop_extent)
    assert (got == O.cell_facet_overlap(cells, V, F, o, vs, np.zeros(3), stop_extent)).all()


def test_classify_cells_box_contacts():
    from pyhexopt import capi
    V, F = mg.cube_surface(n=4, lo=0.0, hi=1.0)
    s = capi.Surface(V, F)
    cells = np.array([[i, j, k, e] for e in (1, 2) for i in range(-3, 6) for j in range(-3, 6) for k in (-1, 0, 1, 4)],
                     dtype=np.int32)
    for mt in (np.zeros(3), np.array([0.25, -0.25, 0.0])):
        got = s.classify_cells(cells, np.zeros(3), 0.25, mt, 0)
        assert (got == O.cell_facet_overlap(cells, V, F, np.zeros(3), 0.25, mt, 0)).all()


def test_subdivision_levels_match():
    """Breadth-first subdivision of one root cell (OctreeGrid::subdivide without grading),
    one predicate batch per level: the same leaf set from GPU and oracle flags."""
    from pyhexopt import capi
    V, F = sphere(3, radius=0.9)
    s = capi.Surface(V, F)
    o, vs = np.array([-1.0, -1.0, -1.0]), 2.0 / 64
    level = np.array([[0, 0, 0, 64]], dtype=np.int32)
    leaves = []
    for _ in range(7):
        g = s.classify_cells(level, o, vs, None, 1)
        r = O.cell_facet_overlap(level, V, F, o, vs, np.zeros(3), 1)
        assert (g == r).all()
        leaves.append(level[g == 0])
        split = level[g == 1]
        if split.size == 0:
            break
        h = split[:, 3:4] // 2
        kids = [np.concatenate([split[:, :3] + h * np.array(c), h], axis=1)
                for c in np.array(mg.HEX_CORNERS)]
        level = np.concatenate(kids).astype(np.int32)
    leaves = np.concatenate(leaves)
    assert (leaves[:, 3] == 1).sum() > 1000


def _voxel_cases():
    cube_V, cube_F, cube_g = cube_aligned_grid()
    g = dict(nx=70, ny=23, nz=70, origin=np.array([-1.25, -1.2, -1.3]), spacing=2.5 / 70)
    nested = merge(*[sphere(2, 0.3 + 0.05 * k) for k in range(17)])   # 34 crossings on the axis
    tV, tF = _torus()
    return {
        "cube_on_vertex_lines": (cube_V, cube_F, cube_g),
        "sphere": (*sphere(3), g),
        "sphere_inward": (*sphere(3, flip=True), g),
        "sphere_doubled": (sphere(3)[0], np.concatenate([sphere(3)[1]] * 2), g),
        "shell": (*merge(sphere(3, 1.0), sphere(3, 0.5, flip=True)), g),
        "nested17": (*nested, g),
        "torus": (tV, tF, g),
        "open_hemisphere": (*_hemisphere(), g),
    }


def _hemisphere():
    V, F = sphere(3)
    keep = V[F].mean(axis=1)[:, 2] < 0   # lower half: -1 crossings, open rim
    return V, F[keep]


@pytest.mark.parametrize("name", list(_voxel_cases().keys()))
def test_voxel_parity_vs_oracle(name):
    from pyhexopt import capi
    V, F, g = _voxel_cases()[name]
    s = capi.Surface(V, F)
    got = s.voxel_parity(g["nx"], g["ny"], g["nz"], g["origin"], g["spacing"])
    ref = O.voxel_parity(V, F, g["nx"], g["ny"], g["nz"], g["origin"], g["spacing"])
    assert got.shape == ref.shape
    bad = np.argwhere(got != ref)
    assert bad.shape[0] == 0, (name, bad[:5].tolist())
    if name != "sphere_inward":
        assert 0 < ref.sum() < ref.size


@pytest.mark.parametrize("name", ["sphere", "shell", "torus", "sphere_doubled", "open_hemisphere"])
def test_cell_parity_vs_oracle(name):
    from pyhexopt import capi
    V, F, _ = _voxel_cases()[name]
    s = capi.Surface(V, F)
    rng = np.random.default_rng(5)
    cells = random_cells(rng, 20000, 64, extents=(1, 2, 4, 8))
    o, sp = np.array([-1.6, -1.6, -1.6]), 3.2 / 64
    got = s.cell_parity(cells, o, sp)
    ref = O.cell_parity(V, F, cells, o, sp)
    assert (got == ref).all()
    assert 0 < ref.sum() < ref.size


def test_hex_inside_vs_oracle():
    from pyhexopt import capi
    SV, SF = _torus()
    V, H = mg.grid_hex(20, -1.5, 1.5)
    V = mg.jitter(V, np.ones(V.shape[0], dtype=bool), 0.01, seed=3)
    h = capi.HexOpt(V, H)
    s = capi.Surface(SV, SF)
    S, flag = h.hex_inside(s, capi.CURRENT)
    P = V[H]
    cen = (P.max(axis=1) + P.min(axis=1)) / 2
    FN, VN, FEN = capi.surface_normals(SV, SF)
    So, _, _, _, _ = O.signed_distance_pseudonormal(cen, SV, SF, FN, VN, FEN)
    assert np.abs(S - So).max() < 1e-14
    assert (flag == (So < 0)).all() and 0 < flag.sum() < flag.size
