"""clean_hex_mesh end to end: the inside test of every hex's box centre on the GPU
(hexopt_hex_inside, the reference's points_inside_mesh), then the host flag passes
(hexclean_clean_flags), against the Python oracle of the passes on the same flags and
against the manifold and one-piece properties of the result.

The grid is coarse against a tilted clipped torus, so its inside flags have spikes, pits
and hexes that meet the rest along an edge or at a vertex; a small second torus beside it
is the piece drop_small_pieces removes.
"""
import numpy as np
import pytest

import hexclean_oracle as HO
from pyhexopt import capi, hexclean, meshgen as mg
from test_hexclean import grid, manifold_and_connected

pytestmark = pytest.mark.gpu


def rot(ax, ay):
    cx, sx, cy, sy = np.cos(ax), np.sin(ax), np.cos(ay), np.sin(ay)
    return np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]]) @ np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])


def scene(ang, n):
    """A tilted clipped torus and a small one beside it (a second piece), and an n^3 grid
    over both."""
    SV0, SF0, _ = mg.clipped_torus_surface(1.0, 0.35, 0.2, nu=96, nband=10, nring=6)
    SV = np.concatenate([SV0 @ rot(ang, 0.4 * ang).T, 0.3 * SV0 + np.array([1.25, 1.25, 0.6])])
    SF = np.concatenate([SF0, SF0 + len(SV0)]).astype(np.int32)
    lo, hi = SV.min(0) - 0.05, SV.max(0) + 0.05
    V, H = grid(n, n, n)
    return SV, SF, lo + V / n * (hi - lo), H


@pytest.mark.parametrize("ang,n", [(0.5, 20), (0.7, 16)])
def test_inside_then_clean_matches_the_oracle(ang, n):
    print("clean_hex_mesh tilted torus", ang, "grid", n, flush=True)
    SV, SF, V, H = scene(ang, n)
    h = capi.HexOpt(V, H)
    s = capi.Surface(SV, SF)
    try:
        S, flag = h.hex_inside(s, capi.CURRENT)
    finally:
        s.close()
        h.close()
    flag = flag.astype(bool)
    assert np.array_equal(flag, S < 0)
    assert 0 < flag.sum() < len(H)
    out = hexclean.clean_flags(len(V), H, flag)
    want = HO.clean_flags(len(V), H.tolist(), flag.tolist())
    assert np.array_equal(out, np.array(want, bool))
    # each pass on its own, in the reference's order, gives the same flags
    a = hexclean.tagging_uneven_element(len(V), H, flag)
    b = hexclean.clean_non_manifold_ve(len(V), H, a)
    c = hexclean.drop_small_pieces(len(V), H, b)
    assert np.array_equal(c, out)
    assert (a != flag).any() and (c != b).any()      # uneven hexes and the small piece go
    manifold_and_connected(len(V), H.tolist(), out.tolist())
