"""GPU octree subdivision (hexopt_octree_build) against the compiled reference's OctreeGrid
(fixtures in tests/golden/octree_cases.npz): the same leaf set for every surface and every
graded/paired mode, the same node count, octree_mesh's hex export consistent with the leaves
and node positions rounded as the reference rounds them; at full size (3.7M leaves) the same
leaf set by SHA-256 and the octree properties of tests/octree_checks.py."""
import os

import numpy as np
import pytest

from octree_checks import check_mesh, check_octree
from pyhexopt import meshgen as mg

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "octree_cases.npz"))


def cases():
    return sorted(tuple(k.split("__")[:2]) for k in FIX.files if k.endswith("__leaves"))


def _sorted(lv):
    return lv[np.lexsort(lv.T[::-1])]


@pytest.mark.parametrize("name,mode", cases())
def test_leaf_set_matches_reference(name, mode):
    from pyhexopt import capi
    print("octree case", name, mode, flush=True)
    V, F = FIX[name + "__V"], FIX[name + "__F"]
    fr = FIX[name + "__frame"]
    origin, vs, mt = fr[:3], float(fr[3]), fr[4:]
    grid = FIX[name + "__grid"]
    s = capi.Surface(V, F)
    out = s.octree(grid, origin, vs, mt, int(FIX[name + "__stop"]), graded=mode[1] == "1", paired=mode[3] == "1")
    ref = FIX["%s__%s__leaves" % (name, mode)]
    assert out["leaves"].shape == ref.shape
    assert np.array_equal(_sorted(out["leaves"]), ref)
    assert len(out["nodes"]) == int(FIX["%s__%s__nnodes" % (name, mode)])
    assert check_mesh(out, mt + origin, vs)


@pytest.mark.parametrize("name", ["sphere64", "torus128"])
def test_targeted_refinement_matches_reference(name):
    """The refinement loop's targeted subdivide (octree.cpp:691-729) of a tenth of the leaves
    at half the STOP extent, on the built octree: the reference's leaf set (SHA-256 of the
    sorted rows) and node count; the refined tree is still graded and paired."""
    import hashlib
    from pyhexopt import capi
    print("octree refine", name, flush=True)
    V, F = FIX[name + "__V"], FIX[name + "__F"]
    fr = FIX[name + "__frame"]
    origin, vs, mt = fr[:3], float(fr[3]), fr[4:]
    grid = FIX[name + "__grid"]
    s = capi.Surface(V, F)
    o = capi.Octree(s, grid, origin, vs, mt, int(FIX[name + "__stop"]))
    try:
        o.refine(s, int(FIX[name + "__refine__stop"]), FIX[name + "__refine__cells"])
        out = o.result()
    finally:
        o.close()
    counts = FIX[name + "__refine__counts"]
    assert len(out["leaves"]) == counts[0] and len(out["nodes"]) == counts[1]
    lv = np.ascontiguousarray(_sorted(out["leaves"]), np.int32)
    assert hashlib.sha256(lv.tobytes()).hexdigest() == str(FIX[name + "__refine__sha256"])
    assert check_octree(out["leaves"], grid)
    assert check_mesh(out, mt + origin, vs)


def test_refine_rejects_cells_off_the_grid():
    from pyhexopt import capi
    V, F = FIX["sphere64__V"], FIX["sphere64__F"]
    fr = FIX["sphere64__frame"]
    s = capi.Surface(V, F)
    o = capi.Octree(s, FIX["sphere64__grid"], fr[:3], float(fr[3]), fr[4:], 2)
    try:
        before = o.result()["leaves"]
        with pytest.raises(RuntimeError, match="not a cell"):
            o.refine(s, 1, np.array([[3, 0, 0, 2]]))      # position not a multiple of the extent
        with pytest.raises(RuntimeError, match="not a cell"):
            o.refine(s, 1, np.array([[0, 0, 0, 3]]))      # extent not a power of two
        o.refine(s, 1, np.zeros((0, 4), np.int32))        # nothing listed: the same tree
        assert np.array_equal(_sorted(o.result()["leaves"]), _sorted(before))
    finally:
        o.close()


def test_compute_octree_composition():
    """voxelization.cpp's compute_octree (:353-391) is hexopt_octree_build without
    mesh_transform at stop extent 1, then compute_sign on the leaves (hexopt_occ_cell_parity):
    the reference's leaf set (SHA-256) and the oracle's inside flags of those leaves."""
    import hashlib
    import oracle_py as O
    from pyhexopt import capi
    V, F = FIX["sphere64__V"], FIX["sphere64__F"]
    fr = FIX["sphere64__frame"]
    origin, vs = fr[:3], float(fr[3])
    s = capi.Surface(V, F)
    out = s.octree(FIX["sphere64__grid"], origin, vs, None, 1)
    counts = FIX["compute_octree__counts"]
    assert len(out["leaves"]) == counts[0] and len(out["nodes"]) == counts[1]
    lv = np.ascontiguousarray(_sorted(out["leaves"]), np.int32)
    assert hashlib.sha256(lv.tobytes()).hexdigest() == str(FIX["compute_octree__sha256"])
    pick = out["leaves"][:: max(1, len(out["leaves"]) // 4000)]
    got = s.cell_parity(pick, origin, vs)
    assert np.array_equal(got, O.cell_parity(V, F, pick, origin, vs))
    assert 0 < got.sum() < len(got)


def test_octree_edge_cases():
    """No split: a surface outside the grid, or a stop extent at the root extent, leaves the
    root cells (and all of them for an anisotropic grid); a single split root with pairing
    splits every root. Leaf and node counts as the reference's OctreeGrid gives them (8 and
    30, 8 and 30, 64 and 135)."""
    from pyhexopt import capi
    from occ_cases import sphere
    V, F = sphere(2, 0.5)
    s = capi.Surface(V, F)
    grid = [64, 32, 16]                 # 4 x 2 x 1 root cells of extent 16
    far = s.octree(grid, np.array([10.0, 10.0, 10.0]), 0.01, None, 1)
    assert len(far["leaves"]) == 8 and set(far["leaves"][:, 3]) == {16}
    assert len(far["nodes"]) == 5 * 3 * 2 and far["levels"] == 1
    stop = s.octree(grid, np.array([-0.6, -0.6, -0.6]), 1.2 / 16, None, 16)
    assert len(stop["leaves"]) == 8
    # the sphere touches only the first root column: pairing splits every root once
    one = s.octree(grid, np.array([-0.6, -0.6, -0.6]), 1.2 / 16, None, 8)
    assert check_octree(one["leaves"], grid)
    assert len(one["leaves"]) >= 8 * 8 and set(one["leaves"][:, 3]) <= {8, 16}
    assert (one["leaves"][:, 3] == 8).sum() == 64 and len(one["leaves"]) == 64
    assert len(one["nodes"]) == 135


def test_octree_mesh_feeds_the_optimiser_context():
    """octree_mesh's export is a valid hex mesh for hexopt_create (a uniform sphere octree)."""
    from pyhexopt import capi
    V, F = FIX["sphere64__V"], FIX["sphere64__F"]
    fr = FIX["sphere64__frame"]
    s = capi.Surface(V, F)
    out = s.octree(FIX["sphere64__grid"], fr[:3], float(fr[3]), fr[4:], 2)
    h = capi.HexOpt(out["V"], out["H"], device=0)
    mn, avg = h.scaled_jacobian()
    assert mn > 0.999 and avg > 0.999   # axis-aligned cubes


def test_full_size_octree_matches_reference():
    """3,717,184 leaves around the 40,000-triangle clipped torus (num_voxels 4096, STOP extent
    8): the SHA-256 of the sorted leaf rows and the node count equal the reference's, and
    tiling, pairing, grading and the export hold."""
    import hashlib
    from pyhexopt import capi
    V, F, _ = mg.clipped_torus_surface(1.0, 0.35, 0.2, nu=200, nband=30, nring=20)
    fr, grid = FIX["torus_full__frame"], FIX["torus_full__grid"]
    origin, vs, mt = fr[:3], float(fr[3]), fr[4:]
    s = capi.Surface(V, F)
    out = s.octree(grid, origin, vs, mt, 8)
    print("full-size octree:", len(out["leaves"]), "leaves", len(out["nodes"]), "nodes", out["levels"], "levels",
          "%.2f ms" % out["ms"], flush=True)
    counts = FIX["torus_full__counts"]
    assert len(out["leaves"]) == counts[0] and len(out["nodes"]) == counts[1]
    lv = np.ascontiguousarray(_sorted(out["leaves"]), np.int32)
    assert hashlib.sha256(lv.tobytes()).hexdigest() == str(FIX["torus_full__sha256"])
    assert check_octree(out["leaves"], grid)
    assert check_mesh(out, mt + origin, vs)
