"""GPU closest point + pseudonormal (hexopt_surface_closest) against igl (compiled-reference
golden vectors), and the deformation loop with on-device tangent-plane projection."""
import numpy as np
import pytest

import oracle_py as O
from helpers import aabb_cases, load, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", aabb_cases())
@pytest.mark.parametrize("given_normals", [True, False])
def test_closest_point_matches_reference(name, given_normals):
    from pyhexopt import capi
    c = load(name)
    nf = c["SF"].shape[0]
    kw = {}
    if given_normals:
        kw = dict(FN=c["ref_FN"].reshape(nf, 3), VN=c["ref_VN"].reshape(-1, 3), FEN=c["ref_FEN"].reshape(nf, 9))
    s = capi.Surface(c["SV"], c["SF"], **kw)
    S, I, C, N = s.closest(c["P"])
    # the tree has igl's shape and is walked in igl's order, and the point-triangle distance
    # sums in Eigen's order: bit for bit, ties included
    assert np.array_equal(I, c["ref_I"])
    assert np.array_equal(C.reshape(-1), c["ref_C"])
    assert np.array_equal(N.reshape(-1), c["ref_N"])
    assert np.array_equal(S, c["ref_S"])


def test_closest_point_vs_oracle_random():
    from pyhexopt import capi, meshgen as mg
    SV, SF = mg.torus_surface(R=1.0, r=0.35, nu=60, nv=30)
    rng = np.random.default_rng(11)
    P = rng.uniform(-1.5, 1.5, size=(4000, 3)) * [1, 1, 0.4]
    FN, VN, FEN = capi.surface_normals(SV, SF)
    S, I, C, N = capi.Surface(SV, SF).closest(P)
    So, Io, Co, No, _ = O.signed_distance_pseudonormal(P, SV, SF, FN, VN, FEN)
    assert np.array_equal(I, Io)
    assert np.array_equal(C, Co) and np.array_equal(N, No) and np.array_equal(S, So)


def test_deformation_loop_matches_reference():
    """grid_hex_meshing.cpp:2945-2959 with one SLIM iteration per outer step: reference
    operator at the current positions, projection of the boundary onto the surface, one
    SLIM iteration, lamda_T schedule."""
    from pyhexopt import capi
    c = load("deform_grid6_superellipsoid")
    n_outer = int(c["n_outer"])
    h = capi.HexOpt(c["V"], c["H"])
    ids = c["idsT"]
    h.set_constraints(idsT=ids, normalT=np.zeros((ids.size, 3)), disT=np.zeros(ids.size), lamT=1.0)
    surf = capi.Surface(c["SV"], c["SF"])
    lam = float(c["lamT0"])
    ref = c["ref_log"].reshape(n_outer, 4)
    Vt = c["ref_Vtrace"].reshape(n_outer, -1, 3)
    for it in range(n_outer):
        h.rest_from_current()
        h.reference()
        h.project_planes(surf, lam)
        E, Eq, Es, log, _ = h.slim_solve(1)
        assert abs(lam - ref[it, 3]) <= 1e-9 * ref[it, 3]
        assert rel([E, Eq], ref[it, :2]) < 1e-9, it
        assert abs(Es - ref[it, 2]) <= 1e-9 * Eq
        assert rel(h.get_positions(), Vt[it]) < 1e-9, it
        lam = min(float(c["lamT0"]) * max(Eq / Es * lam, 1.0), float(c["lamT_bound"]))
