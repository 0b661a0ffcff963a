"""GPU octree subdivision (hexopt_octree_build) against the compiled reference's OctreeGrid
(fixtures in tests/golden/octree_cases.npz): the same leaf set for every surface and every
graded/paired mode, the same node count, octree_mesh's hex export consistent with the leaves
and node positions rounded as the reference rounds them; at full size (about a million
leaves) the octree properties of tests/octree_checks.py."""
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


def test_full_size_octree_properties():
    """About a million leaves around the 40,000-triangle clipped torus: tiling, pairing and
    grading hold, and the export is consistent."""
    from pyhexopt import capi
    V, F, _ = mg.clipped_torus_surface(1.0, 0.35, 0.2, nu=200, nband=30, nring=20)
    lo, hi = V.min(0), V.max(0)
    num_voxels = 1 << 12
    vs = float((hi - lo).max()) / num_voxels
    grid = [1 << int(np.ceil(np.log2(np.ceil(e / vs)))) for e in (hi - lo)]
    origin = lo
    mt = (lo + hi) / 2 - (origin + (origin + vs * np.array(grid, np.float64))) * 0.5
    s = capi.Surface(V, F)
    out = s.octree(grid, origin, vs, mt, 8)
    print("full-size octree:", len(out["leaves"]), "leaves", len(out["nodes"]), "nodes", out["levels"], "levels",
          "%.2f ms" % out["ms"], flush=True)
    assert len(out["leaves"]) > 200000
    assert check_octree(out["leaves"], grid)
    assert check_mesh(out, mt + origin, vs)
