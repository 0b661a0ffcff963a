"""Host-side normals of the input surface against igl's (golden vectors of the compiled
reference: per_face_normals, per_vertex_normals ANGLE, per_edge_normals UNIFORM)."""
import numpy as np
import pytest

from helpers import aabb_cases, load
from pyhexopt import capi


@pytest.mark.parametrize("name", aabb_cases())
def test_normals_match_reference(name):
    c = load(name)
    nf = c["SF"].shape[0]
    FN, VN, FEN = capi.surface_normals(c["SV"], c["SF"])
    assert np.abs(FN - c["ref_FN"].reshape(nf, 3)).max() < 1e-15
    assert np.abs(VN - c["ref_VN"].reshape(-1, 3)).max() < 1e-14
    assert np.abs(FEN - c["ref_FEN"].reshape(nf, 9)).max() < 1e-15
