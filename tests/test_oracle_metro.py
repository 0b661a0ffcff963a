"""The metro Hausdorff oracle (vertex sampling) against the compiled reference's
compute(const Mesh&, const Mesh&, ...) on committed fixtures, and against closed forms:
concentric spheres (every distance is the radius gap up to facet sagitta) and a mesh
measured against itself (zero)."""
import numpy as np
import pytest

import oracle_py as O
from helpers import load

CASES = ["metro_sphere_superellipsoid", "metro_unreferenced_vertex", "metro_torus_hexboundary"]


@pytest.mark.parametrize("name", CASES)
def test_metro_matches_reference(name):
    c = load(name)
    D, mx, mean = O.metro_hausdorff(c["V0"], c["F0"], c["V1"], c["F1"])
    assert D == c["ref_bbox_diagonal"][0]
    assert mx == c["ref_max"][0]
    assert mean == c["ref_mean"][0]


def test_metro_closed_forms():
    from pyhexopt import meshgen as mg
    V, F = mg.icosphere(2)
    D, mx, mean = O.metro_hausdorff(V, F, V, F)
    assert mx == 0 and mean == 0
    # grown box: [-1, 1]^3 plus 2% of its diagonal d0 on each side of every axis
    d0 = 2 * np.sqrt(3)
    assert abs(D - np.sqrt(3) * (2 + 2 * 0.02 * d0)) < 1e-12
    V2 = V * 1.5
    D, mx, mean = O.metro_hausdorff(V, F, V2, F)
    # vertices of the outer sphere are 0.5 from the inner sphere's vertices along rays; the
    # inner polyhedron's faces lie slightly inside radius 1, so distances are in [0.5, 0.5 + sag]
    assert 0.5 - 1e-12 <= mean <= mx <= 0.5 + 0.05
