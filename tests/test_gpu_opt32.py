"""FP32+adjoint optimiser (hexopt_opt32_solve) against the FP64 SLIM iteration
(hexopt_slim_solve) on the same inputs.

Both take the same step. Eigen's CG warm-started at x on L x' = rhs, and the FP32 CG from 0
on L d = rhs - L x = -dE/dx / 2, build the same Krylov iterates (x' = x + d). So one
iteration agrees to FP32 accuracy when both line searches accept the same step, and over a
few iterations the energies track. Also covered: bitwise run-to-run determinism of the
fused CG (its reductions are in a fixed order), many-tile meshes, feature-line rows, and a
vertex that belongs to no hex."""
import numpy as np
import pytest

from helpers import gpu_constraint_kwargs, load

pytestmark = pytest.mark.gpu

CASES = ["slim_grid4_sd", "slim_grid5_inverted", "slim_grid6_features"]


def make(c, V=None, H=None, kw=None):
    from pyhexopt import capi
    h = capi.HexOpt(c["V"] if V is None else V, c["H"] if H is None else H)
    h.reference(hflag=c.get("hflag"), cube=True)
    h.set_constraints(**(gpu_constraint_kwargs(c) if kw is None else kw))
    return h


def _compare_first_iteration(h64, h32, wo, disp_tol):
    X0 = h64.get_positions()
    E64, _, _, log64, _ = h64.slim_solve(1, weight_opt=wo)
    E32, log32 = h32.opt32_solve(1, weight_opt=wo)
    X64, X32 = h64.get_positions(), h32.get_positions()
    assert log32[0, 4] == pytest.approx(log64[0, 4], rel=1e-3)     # accepted step
    disp = np.abs(X64 - X0).max()
    assert disp > 0
    assert np.abs(X32 - X64).max() < disp_tol * disp
    assert abs(E32 - E64) < 1e-5 * abs(E64)
    return log64, log32


@pytest.mark.parametrize("name", CASES)
def test_first_iteration_matches_fp64(name):
    c = load(name)
    wo = float(c["weight_opt"])
    log64, log32 = _compare_first_iteration(make(c), make(c), wo, 2e-3)
    assert int(log32[0, 5]) == int(log64[0, 5])   # CG iterations


def test_many_tiles_tracks_fp64():
    from pyhexopt import meshgen as mg
    n = 20
    V, H = mg.grid_hex(n)
    V = mg.jitter(V, ~mg.grid_boundary_mask(n), 0.15 * 2.0 / n, seed=6)
    V, _ = mg.invert_hexes(V, H, 0.02, seed=6)
    c = {"V": V, "H": H}
    h64, h32 = make(c, kw={}), make(c, kw={})
    assert h32.layout_info()["tiles"] > 20
    _compare_first_iteration(h64, h32, 1.0, 2e-3)
    E64, _, _, _, _ = h64.slim_solve(3)
    E32, _ = h32.opt32_solve(3)
    assert abs(E32 - E64) < 1e-4 * abs(E64)


def test_opt32_bitwise_deterministic():
    c = load("slim_grid6_features")
    wo = float(c["weight_opt"])
    a, b = make(c), make(c)
    a.opt32_solve(3, weight_opt=wo)
    b.opt32_solve(3, weight_opt=wo)
    assert np.array_equal(a.get_positions(), b.get_positions())


def test_vertex_without_hexes():
    """An extra vertex outside every hex, pulled by a corner constraint: its rows of L are
    prox + lamC only, handled by the reducing block of the fused CG."""
    c = load("slim_grid4_sd")
    V = np.concatenate([c["V"], [[2.0, 2.0, 2.0]]])
    kw = dict(idsC=np.array([V.shape[0] - 1], dtype=np.int32), Cpos=np.array([[2.5, 2.0, 1.5]]), lamC=10.0)
    wo = float(c["weight_opt"])
    _compare_first_iteration(make(c, V=V, kw=kw), make(c, V=V, kw=kw), wo, 2e-3)
