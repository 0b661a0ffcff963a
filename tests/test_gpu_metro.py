"""GPU metro Hausdorff distance (hexopt_metro_hausdorff) against the compiled reference's
compute(const Mesh&, const Mesh&, ...) fixtures and the oracle: the bounding-box diagonal,
the maximum and the mean equal to the last bit (vcg's edge-plane point-to-face distance, the
samples added in vertex order), on the fixture cases and on the full-size bench case (the
clipped torus against its 1M-hex mesh's boundary)."""
import os

import numpy as np
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
    assert mx == c["ref_max"][0]
    assert mean == c["ref_mean"][0]
    Do, mxo, meano = O.metro_hausdorff(c["V0"], c["F0"], c["V1"], c["F1"])
    assert (D, mx, mean) == (Do, mxo, meano)


def test_metro_full_size_matches_reference():
    from pyhexopt import capi, meshgen as mg
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "metro_full.npz"))
    V, H, _ = mg.torus_hex_mesh(int(z["target"]), 1.0, 0.35, 0.2)
    assert H.shape[0] == int(z["hexes"])
    T = mg.hex_boundary_triangles(H).astype(np.int32)
    SV, SF, _ = mg.clipped_torus_surface(1.0, 0.35, 0.2, nu=200, nband=30, nring=20)
    assert len(T) == int(z["hex_boundary_triangles"])
    D, mx, mean = capi.metro_hausdorff(SV, SF, V, T)
    print("full-size metro", D, mx, mean, flush=True)
    assert (D, mx, mean) == (float(z["ref_bbox_diagonal"]), float(z["ref_max"]), float(z["ref_mean"]))
