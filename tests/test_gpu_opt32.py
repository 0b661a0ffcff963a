"""FP32+adjoint optimiser (hexopt_opt32_solve) against the FP64 SLIM iteration
(hexopt_slim_solve) on the same inputs.

Both take the same step in exact arithmetic. Eigen's CG warm-started at x on L x' = rhs,
and the FP32 CG from 0 on L d = rhs - L x = -dE/dx / 2, build the same Krylov iterates
(x' = x + d). After the reference's 20 CG steps, though, the iterate of these
ill-conditioned systems moves with rounding along directions the energy barely sees (the
compiled reference's own reordering sensitivity reaches 1e-4 in FP64). So the FP32
iteration must reach the same energy and step, with a CG residual of the same size, rather
than the same positions. Only the well-conditioned case also pins positions. The FP32 CG's
recurrence residual must equal the residual recomputed in FP64 (the fused reductions are
right). Also covered: bitwise run-to-run determinism, many-tile meshes, feature-line rows,
and a vertex that belongs to no hex."""
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


def _compare_first_iteration(h64, h32, wo, disp_tol=None):
    X0 = h64.get_positions()
    E0 = sum(h64.energy(weight_opt=wo)) / (wo * 8 * h64.nH)
    E64, _, _, log64, _ = h64.slim_solve(1, weight_opt=wo)
    E32, log32 = h32.opt32_solve(1, weight_opt=wo)
    X64, X32 = h64.get_positions(), h32.get_positions()
    assert log32[0, 4] == pytest.approx(log64[0, 4], rel=1e-3)     # accepted step
    assert int(log32[0, 5]) == int(log64[0, 5])                    # CG iterations
    # same energy: to 5e-5, or to 1% of the iteration's decrease on strongly inverted meshes
    assert abs(E32 - E64) < max(5e-5 * abs(E64), 1e-2 * (E0 - E64))
    # the FP32 CG's recurrence residual is the true residual, recomputed in FP64 against the
    # FP64 adjoint's right-hand side (which differs from the FP32 one by FP32 rounding) ...
    assert abs(log32[0, 7] - log32[0, 6]) <= 1e-2 * log32[0, 6] + 3e-5
    # ... and of the size the FP64 CG reaches in as many steps
    assert log32[0, 7] < 5 * log64[0, 6] + 1e-6
    if disp_tol is not None:
        disp = np.abs(X64 - X0).max()
        assert disp > 0 and np.abs(X32 - X64).max() < disp_tol * disp
    return log64, log32


@pytest.mark.parametrize("name", CASES)
def test_first_iteration_matches_fp64(name):
    c = load(name)
    wo = float(c["weight_opt"])
    _compare_first_iteration(make(c), make(c), wo, 1e-4 if name == "slim_grid4_sd" else None)


def test_many_tiles_tracks_fp64():
    from pyhexopt import meshgen as mg
    n = 20
    V, H = mg.grid_hex(n)
    V = mg.jitter(V, ~mg.grid_boundary_mask(n), 0.15 * 2.0 / n, seed=6)
    V, _ = mg.invert_hexes(V, H, 0.02, seed=6)
    c = {"V": V, "H": H}
    h64, h32 = make(c, kw={}), make(c, kw={})
    assert h32.layout_info()["tiles"] > 20
    E0 = sum(h64.energy()) / (8 * H.shape[0])
    _compare_first_iteration(h64, h32, 1.0)
    E64, _, _, _, _ = h64.slim_solve(3)
    E32, _ = h32.opt32_solve(3)
    assert E32 < E0 and E64 < E0 and abs(E32 - E64) < 2e-2 * (E0 - E64)


def test_colour_scatter_tiles_track_fp64():
    """Half the hexes numbered from rotated corners, so most tiles of the fused FP32 CG
    operator scatter by hex colour instead of by corner index."""
    from pyhexopt import meshgen as mg
    n = 14
    V, H = mg.grid_hex(n)
    V = mg.jitter(V, ~mg.grid_boundary_mask(n), 0.15 * 2.0 / n, seed=8)
    H = mg.rotate_hex_corners(H, 0.5, seed=8)
    c = {"V": V, "H": H}
    h64, h32 = make(c, kw={}), make(c, kw={})
    info = h32.layout_info()
    assert info["corner_tiles"] < info["tiles"] // 2
    _compare_first_iteration(h64, h32, 1.0)


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
    _compare_first_iteration(make(c, V=V, kw=kw), make(c, V=V, kw=kw), wo, 1e-4)


def _direction(h, X0, log):
    """The iteration's direction d from x' = x + step d (vertex rows, flattened [x y z])."""
    step = float(log[0, 4])
    assert step > 0
    return ((h.get_positions() - X0) / step).T.reshape(-1)


def test_refinement_moves_toward_fp64_solve():
    """Iterative refinement: the residual against the FP64 right-hand side shrinks after the
    correction, and the direction moves toward the converged solve of the FP64 system
    (formed by hexopt_debug_system with the FP64 weights, solved here by CG to 1e-12)."""
    from scipy.sparse.linalg import LinearOperator, cg
    c = load("slim_grid6_features")
    wo = float(c["weight_opt"])
    h64, h0, h1 = make(c), make(c), make(c)
    X0 = h64.get_positions()
    nV = X0.shape[0]
    nL = int(c["idsL"].size)
    n = 3 * nV + nL
    x0 = np.concatenate([X0.T.reshape(-1), np.zeros(nL)])
    rhs, invd, Lx = h64.debug_system(x0, weight_opt=wo)
    b = rhs - Lx                                      # the direction system's right-hand side
    # the FP64 adjoint gives the same right-hand side: -dE/dx / 2 at a_L = 0
    g = h64.energy_grad(X0, weight_opt=wo, aL=np.zeros(nL))
    assert np.abs(-0.5 * g.T.reshape(-1) - b[:3 * nV]).max() < 1e-9 * np.abs(b).max()
    op = LinearOperator((n, n), matvec=lambda p: h64.debug_system(p, weight_opt=wo)[2], dtype=np.float64)
    dstar, info = cg(op, b, rtol=1e-12, atol=0.0, maxiter=5000, M=LinearOperator((n, n), matvec=lambda r: invd * r))
    assert info == 0
    _, log0 = h0.opt32_solve(1, weight_opt=wo, refine_iters=0, residual_every=1)
    _, log1 = h1.opt32_solve(1, weight_opt=wo, refine_iters=10, residual_every=1)
    assert log0[0, 9] == -1.0 and log1[0, 9] >= 0.0
    d0, d1 = _direction(h0, X0, log0), _direction(h1, X0, log1)
    ds = dstar[:3 * nV]
    e0, e1 = np.linalg.norm(d0 - ds), np.linalg.norm(d1 - ds)
    print("refinement: residual %.3e -> %.3e, |d - d*| / |d*| %.3e -> %.3e" %
          (log1[0, 7], log1[0, 9], e0 / np.linalg.norm(ds), e1 / np.linalg.norm(ds)), flush=True)
    # same starting residual (same FP32 solve), smaller after the correction: a restarted
    # Jacobi-CG of 10 steps removes part of what the 20-step solve left
    assert log1[0, 7] == pytest.approx(log0[0, 7], rel=1e-6)
    assert log1[0, 9] < 0.9 * log1[0, 7]
    assert e1 < 0.9 * e0
