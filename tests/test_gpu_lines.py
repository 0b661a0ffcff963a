"""GPU feature-curve projection (hexopt_project_lines) against the oracle, bit for bit:
LINE vertices of a jittered grid projected onto the four rim circles of the clipped torus
and onto open polylines, including exact ties at shared curve vertices."""
import numpy as np
import pytest

import oracle_py as O
from pyhexopt import meshgen as mg

pytestmark = pytest.mark.gpu


def _curves():
    R, r, zc = 1.0, 0.35, 0.2
    w = np.sqrt(r * r - zc * zc)
    V, curves, circles = [], [], []
    n = 200
    u = np.arange(n) * 2 * np.pi / n
    for rad in (R - w, R + w):
        for z in (-zc, zc):
            base = len(V)
            V += [[rad * np.cos(t), rad * np.sin(t), z] for t in u]
            curves.append(list(range(base, base + n)))
            circles.append(1)
    # open polylines: a zig-zag (ties at its corners) and a straight one with short segments
    base = len(V)
    V += [[-1.0 + 0.25 * k, 0.3 * (k % 2), 0.0] for k in range(9)]
    curves.append(list(range(base, base + 9)))
    circles.append(0)
    base = len(V)
    V += [[0.0, -1.0 + 1e-4 * k, 0.1] for k in range(5)] + [[0.0, 1.0, 0.1]]
    curves.append(list(range(base, base + 6)))
    circles.append(0)
    base = len(V)
    V += [[0.3, 0.3, 0.3]]
    curves.append([base])
    circles.append(0)
    return np.array(V), curves, np.array(circles, dtype=np.uint8)


def test_project_lines_vs_oracle():
    from pyhexopt import capi
    SV, curves, circles = _curves()
    V, H = mg.grid_hex(16, -1.5, 1.5)
    V = mg.jitter(V, np.ones(V.shape[0], dtype=bool), 0.02, seed=9)
    h = capi.HexOpt(V, H)
    cv = capi.Curves(SV, curves, circles)
    rng = np.random.default_rng(1)
    ids = rng.choice(V.shape[0], size=3000, replace=False).astype(np.int32)
    cid = rng.integers(0, len(curves) - 1, size=ids.size).astype(np.int32)   # not the one-vertex curve
    o, a = h.project_lines(cv, ids, cid)
    oo, ao = O.line_projection(SV, curves, circles, V[ids], cid)
    assert (o == oo).all() and (a == ao).all()
    # exact ties: points placed beyond each zig-zag corner, equidistant from both segments
    zz = np.array(curves[4])
    corner = SV[zz[1:-1]]
    for k, c in enumerate(corner):
        V[k] = c + np.array([0.0, 0.5 if k % 2 == 0 else -0.5, 0.0])
    h2 = capi.HexOpt(V, H)
    ids2 = np.arange(corner.shape[0], dtype=np.int32)
    o2, a2 = h2.project_lines(cv, ids2, np.full(ids2.size, 4, dtype=np.int32))
    oo2, ao2 = O.line_projection(SV, curves, circles, V[ids2], np.full(ids2.size, 4))
    assert (o2 == oo2).all() and (a2 == ao2).all()
    # the one-vertex curve: the vertex itself and a NaN tangent, as the reference computes
    o3, a3 = h.project_lines(cv, ids[:5], np.full(5, len(curves) - 1, dtype=np.int32))
    assert (o3 == SV[-1]).all() and np.isnan(a3).all()


def test_refresh_lines_equals_set_constraints():
    """hexopt_refresh_lines (projection and re-folding on the device) gives bit for bit the
    system and the SLIM iteration that hexopt_set_constraints gives with the same projected
    origins and axes."""
    from pyhexopt import capi
    SV, curves, circles = _curves()
    n = 8
    V, H = mg.grid_hex(n, -1.5, 1.5)
    V = mg.jitter(V, ~mg.grid_boundary_mask(n), 0.02, seed=4)
    cv = capi.Curves(SV, curves, circles)
    rng = np.random.default_rng(3)
    ids = np.sort(rng.choice(V.shape[0], size=60, replace=False)).astype(np.int32)
    cid = rng.integers(0, 6, size=ids.size).astype(np.int32)
    lam = 10.0
    h1 = capi.HexOpt(V, H)
    h1.reference()
    h1.set_constraints(idsL=ids, axaL=np.zeros((ids.size, 3)), originL=np.zeros((ids.size, 3)), lamL=1.0)
    h1.refresh_lines(cv, cid, lam)
    o, a = h1.project_lines(cv, ids, cid)
    h2 = capi.HexOpt(V, H)
    h2.reference()
    h2.set_constraints(idsL=ids, axaL=a, originL=o, lamL=lam)
    p = np.random.default_rng(0).normal(size=3 * V.shape[0] + ids.size)
    for x, y in zip(h1.debug_system(p), h2.debug_system(p)):
        assert np.array_equal(x, y)
    r1, r2 = h1.slim_solve(2), h2.slim_solve(2)
    assert r1[0] == r2[0] and np.array_equal(h1.get_positions(), h2.get_positions())
