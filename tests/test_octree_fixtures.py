"""The octree fixtures (tests/golden/octree_cases.npz, written by the compiled reference's
OctreeGrid from grid_meshing/octree.cpp) and the property checks the GPU octree tests use
(CPU only).

The reference's own verdicts (is2to1Graded, isPaired) agree with the numpy properties of
tests/octree_checks.py on every fixture; planted mistakes (a split leaf, the ungraded leaf set
checked as graded) fail those properties, so they are not vacuous; where the driver is built,
it reproduces the fixture.
"""
import os
import sys

import numpy as np
import pytest

from octree_checks import check_octree

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import octcases as oc  # noqa: E402

FIX = np.load(os.path.join(HERE, "golden", "octree_cases.npz"))
live = pytest.mark.skipif(not oc.available(), reason="oracle/_ref/octree_ref not built (reference tree absent)")


def cases():
    out = []
    for k in FIX.files:
        if k.endswith("__leaves"):
            name, mode, _ = k.split("__")
            out.append((name, mode))
    return sorted(out)


@pytest.mark.parametrize("name,mode", cases())
def test_reference_leaf_sets_have_the_octree_properties(name, mode):
    lv = FIX["%s__%s__leaves" % (name, mode)]
    g, p = int(mode[1]), int(mode[3])
    flags = FIX["%s__%s__flags" % (name, mode)]
    # the reference's own checks report what was asked for
    if g:
        assert flags[0] == 1
    if p:
        assert flags[1] == 1
    assert check_octree(lv, FIX[name + "__grid"], graded=bool(g), paired=bool(p))


@pytest.mark.parametrize("name", ["sphere64", "torus128"])
def test_reference_refinement_stays_graded_and_paired(name):
    assert list(FIX[name + "__refine__flags"]) == [1, 1]
    assert FIX[name + "__refine__counts"][0] > len(FIX[name + "__g1p1__leaves"])


@live
def test_driver_reproduces_refinement_fixture():
    import hashlib
    name = "sphere64"
    V, F = FIX[name + "__V"], FIX[name + "__F"]
    fr = FIX[name + "__frame"]
    r = oc.ref_octree(V, F, FIX[name + "__grid"].tolist(), fr[:3], fr[3], fr[4:], int(FIX[name + "__stop"]),
                      refine=(int(FIX[name + "__refine__stop"]), FIX[name + "__refine__cells"]))
    lv = np.ascontiguousarray(r["leaves"][np.lexsort(r["leaves"].T[::-1])], np.int32)
    assert hashlib.sha256(lv.tobytes()).hexdigest() == str(FIX[name + "__refine__sha256"])


def test_planted_split_breaks_pairing():
    lv = FIX["sphere64__g1p1__leaves"].astype(np.int64)
    grid = FIX["sphere64__grid"]
    i = int(np.nonzero((lv[:, 3] > 1) & (lv[:, 3] < grid.min()))[0][0])
    c = lv[i]
    h = c[3] // 2
    kids = np.array([[c[0] + h * a, c[1] + h * b, c[2] + h * d, h] for d in (0, 1) for b in (0, 1) for a in (0, 1)])
    bad = np.concatenate([np.delete(lv, i, axis=0), kids])
    check_octree(bad, grid, graded=False, paired=False)   # still a tiling
    with pytest.raises(AssertionError, match="paired"):
        check_octree(bad, grid, graded=False, paired=True)


def test_ungraded_set_fails_the_grading_property():
    lv = FIX["sphere64__g0p1__leaves"]
    grid = FIX["sphere64__grid"]
    assert len(lv) != len(FIX["sphere64__g1p1__leaves"])
    with pytest.raises(AssertionError, match="graded"):
        check_octree(lv, grid, graded=True, paired=True)


def test_planted_overlap_and_gap_break_the_tiling():
    lv = FIX["torus128__g1p1__leaves"].astype(np.int64)
    grid = FIX["torus128__grid"]
    with pytest.raises(AssertionError, match="tile"):
        check_octree(lv[1:], grid, graded=False, paired=False)
    dup = np.concatenate([lv, lv[:1]])
    with pytest.raises(AssertionError):
        check_octree(dup, grid, graded=False, paired=False)


@live
def test_driver_reproduces_fixture():
    name = "ellipsoid64"
    V, F = FIX[name + "__V"], FIX[name + "__F"]
    fr = FIX[name + "__frame"]
    r = oc.ref_octree(V, F, FIX[name + "__grid"].tolist(), fr[:3], fr[3], fr[4:], int(FIX[name + "__stop"]))
    lv = r["leaves"]
    assert np.array_equal(lv[np.lexsort(lv.T[::-1])], FIX[name + "__g1p1__leaves"])
    assert len(r["nodes"]) == int(FIX[name + "__g1p1__nnodes"])
