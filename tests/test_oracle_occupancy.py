"""The occupancy oracle (subdivision predicate, ray-parity inside tests) against closed forms:
exact inside/outside of a cube sampled on its own vertex, edge and diagonal lines (where the
SOS tie-break must count every crossing exactly once), spheres away from their surface,
orientation and multiplicity rules of the signed crossing count, and box contact cases of
the closed-interval predicate."""
import numpy as np
import pytest

import oracle_py as O
from occ_cases import cell_centres, cube_aligned_grid, merge, random_cells, sphere, voxel_centres


def test_voxel_parity_cube_on_vertex_lines():
    V, F, g = cube_aligned_grid()
    inside = O.voxel_parity(V, F, g["nx"], g["ny"], g["nz"], g["origin"], g["spacing"])
    P = voxel_centres(g["nx"], g["ny"], g["nz"], g["origin"], g["spacing"])
    x, y, z = P[..., 0], P[..., 1], P[..., 2]
    strict_in = (x > 0) & (x < 1) & (y > 0) & (y < 1) & (z > 0) & (z < 1)
    strict_out = (x < 0) | (x > 1) | (y < 0) | (y > 1) | (z < 0) | (z > 1)
    # columns through interior vertices (x, y in {0.25, 0.5, 0.75}), edges and diagonals
    assert (inside[strict_in] == 1).all() and strict_in.sum() == 7 * 7 * 8
    assert (inside[strict_out] == 0).all()


@pytest.mark.parametrize("level", [2, 3])
def test_voxel_parity_sphere(level):
    V, F = sphere(level, radius=1.0)
    g = dict(nx=17, ny=15, nz=19, origin=np.array([-1.3, -1.2, -1.4]), spacing=0.15)
    inside = O.voxel_parity(V, F, **g)
    r = np.linalg.norm(voxel_centres(**g), axis=-1)
    sag = 0.05 if level == 2 else 0.015   # facet sagitta of the inscribed icosphere
    assert (inside[r < 1 - sag] == 1).all() and (inside[r > 1] == 0).all()


def test_voxel_parity_orientation_and_multiplicity():
    g = dict(nx=12, ny=12, nz=12, origin=np.array([-1.2, -1.2, -1.2]), spacing=0.2)
    r = np.linalg.norm(voxel_centres(**g), axis=-1)
    single = O.voxel_parity(*sphere(2), **g)
    # inward-facing surface: crossings below an inner point sum to +1, never inside
    assert (O.voxel_parity(*sphere(2, flip=True), **g) == 0).all()
    # every facet twice: the sum is -2 inside, still negative
    V, F = sphere(2)
    assert (O.voxel_parity(V, np.concatenate([F, F]), **g) == single).all()
    # hollow shell: an inward inner sphere cancels the outer one inside it
    shell = O.voxel_parity(*merge(sphere(2, 1.0), sphere(2, 0.5, flip=True)), **g)
    assert (shell[(r < 0.45)] == 0).all() and (shell[(r > 0.55) & (r < 0.95)] == 1).all()
    # nested outward spheres: -2 inside the inner one, still inside
    nested = O.voxel_parity(*merge(sphere(2, 1.0), sphere(2, 0.5)), **g)
    assert (nested[r < 0.45] == 1).all()


def test_voxel_parity_open_triangle():
    # one triangle at z = 0.5; its xy orientation decides the crossing sign
    V = np.array([[0.0, 0.0, 0.5], [1.0, 0.0, 0.5], [0.0, 1.0, 0.5]])
    g = dict(nx=4, ny=4, nz=4, origin=np.array([0.0, 0.0, 0.0]), spacing=0.25)
    up = O.voxel_parity(V, np.array([[0, 1, 2]], dtype=np.int32), **g)
    down = O.voxel_parity(V, np.array([[0, 2, 1]], dtype=np.int32), **g)
    P = voxel_centres(**g)
    foot = (P[..., 0] + P[..., 1] < 1.0)
    above = P[..., 2] > 0.5
    assert (up == 0).all()                       # +1 below: never negative
    assert (down[foot & above] == 1).all() and (down[~(foot & above)] == 0).all()


# Queries on the boundary of the counter-clockwise triangle (0,0), (1,0), (0,1): which ones
# the SOS orientation (voxelization.cpp:62-72) counts, worked out by hand from its rules.
# A consistent tie-break counts a point on an edge shared by two triangles in exactly one of
# them; these cases also fix WHICH one, i.e. the reference's choice of perturbation.
SOS_CASES = [((0.5, 0.0), True), ((0.0, 0.5), True), ((0.0, 0.0), True),
             ((0.5, 0.5), False), ((1.0, 0.0), False), ((0.0, 1.0), False)]


@pytest.mark.parametrize("q,counted", SOS_CASES)
def test_voxel_parity_sos_tie_break(q, counted):
    V = np.array([[0.0, 0.0, 0.5], [1.0, 0.0, 0.5], [0.0, 1.0, 0.5],
                  [-5.0, -5.0, 0.25], [15.0, -5.0, 0.25], [-5.0, 15.0, 0.25]])
    F = np.array([[0, 1, 2], [3, 5, 4]], dtype=np.int32)   # +1 crossing over a -1 floor
    # a single voxel whose centre is (q, 0.75): (0 + 0.5) * 1 + (q - 0.5) is exact here
    origin = np.array([q[0] - 0.5, q[1] - 0.5, 0.25])
    inside = O.voxel_parity(V, F, 1, 1, 1, origin, 1.0)
    assert inside.item() == (0 if counted else 1)


def test_cell_parity_sphere():
    rng = np.random.default_rng(4)
    V, F = sphere(3, radius=1.0)
    origin = np.array([-1.6, -1.6, -1.6])
    spacing = 3.2 / 64
    cells = random_cells(rng, 3000, 64)
    inside = O.cell_parity(V, F, cells, origin, spacing)
    r = np.linalg.norm(cell_centres(cells, origin, spacing), axis=1)
    assert (inside[r < 0.985] == 1).all() and (inside[r > 1] == 0).all()
    # the reduced-list rule: a doubled surface still reads as inside once
    assert (O.cell_parity(V, np.concatenate([F, F]), cells, origin, spacing) == inside).all()
    shell = O.cell_parity(*merge(sphere(3, 1.0), sphere(3, 0.5, flip=True)), cells, origin, spacing)
    assert (shell[r < 0.48] == 0).all() and (shell[(r > 0.51) & (r < 0.985)] == 1).all()


def test_cell_facet_overlap_contacts():
    from pyhexopt import meshgen as mg
    V, F = mg.cube_surface(n=4, lo=0.0, hi=1.0)
    o, mt = np.zeros(3), np.zeros(3)
    cells = np.array([[0, 0, 0, 1],      # touches three sides
                      [1, 1, 1, 1],      # interior, clear of every facet box
                      [1, 1, 1, 2],      # interior [0.25, 0.75]^3
                      [0, 0, 0, 4],      # the whole cube
                      [-1, 1, 1, 1],     # outside, its max x face lies on x = 0
                      [-2, 1, 1, 1],     # outside, a gap of 0.25
                      [4, 1, 1, 1],      # outside, its min x face lies on x = 1
                      [0, 0, 0, 1]], dtype=np.int32)
    flag = O.cell_facet_overlap(cells, V, F, o, 0.25, mt, 0)
    assert flag.tolist() == [1, 0, 0, 1, 1, 0, 1, 1]
    # stop_extent: cells no larger than it are never split
    flag = O.cell_facet_overlap(cells, V, F, o, 0.25, mt, 1)
    assert flag.tolist() == [0, 0, 0, 1, 0, 0, 0, 0]
    # mesh_transform shifts the grid: +0.25 in x moves cell (-2, ...) onto x = 0
    flag = O.cell_facet_overlap(cells, V, F, o, 0.25, np.array([0.25, 0, 0]), 0)
    assert flag[5] == 1 and flag[4] == 1


def test_cell_facet_overlap_random_vs_box_test():
    rng = np.random.default_rng(2)
    V, F = sphere(2, radius=0.7)
    cells = random_cells(rng, 2000, 32)
    o, vs, mt = np.array([-1.0, -1.0, -1.0]), 2.0 / 32, np.array([0.01, -0.02, 0.0])
    flag = O.cell_facet_overlap(cells, V, F, o, vs, mt, 1)
    fmin, fmax = V[F].min(axis=1), V[F].max(axis=1)
    c = cells.astype(np.float64)
    mn = (mt + o)[None, :] + vs * c[:, :3]
    mx = mn + vs * c[:, 3:4]
    ov = ((mx[:, None, :] >= fmin[None]) & (mn[:, None, :] <= fmax[None])).all(axis=2).any(axis=1)
    assert ((cells[:, 3] > 1) & ov).astype(np.uint8).tolist() == flag.tolist()
    assert 0 < flag.sum() < flag.size
