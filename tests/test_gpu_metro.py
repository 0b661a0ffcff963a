"""GPU metro Hausdorff distance (hexopt_metro_hausdorff) against the compiled reference's
compute(const Mesh&, const Mesh&, ...) fixtures and the oracle."""
import pytest

import oracle_py as O
from helpers import load

pytestmark = pytest.mark.gpu

CASES = ["metro_sphere_superellipsoid", "metro_unreferenced_vertex", "metro_torus_hexboundary"]


@pytest.mark.parametrize("name", CASES)
def test_metro_matches_reference(name):
    from pyhexopt import capi
    c = load(name)
    D, mx, mean = capi.metro_hausdorff(c["V0"], c["F0"], c["V1"], c["F1"])
    assert D == c["ref_bbox_diagonal"][0]
    assert abs(mx - c["ref_max"][0]) <= 1e-15 * c["ref_max"][0]
    # the mean is a sum over thousands of samples: its order differs from the reference's
    assert abs(mean - c["ref_mean"][0]) <= 1e-13 * c["ref_mean"][0]
    Do, mxo, meano = O.metro_hausdorff(c["V0"], c["F0"], c["V1"], c["F1"])
    assert mx == mxo and abs(mean - meano) <= 1e-13 * meano
