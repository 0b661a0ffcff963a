"""The feature-curve projection oracle (LINE branch of dirty_graph_projection with
point_line_projection) against closed forms: perpendicular feet, clamping at curve ends,
the lower segment index winning when two segments give the same point, the closing segment
of a circle, segments shorter than sqrt(1e-7) projecting to their first vertex, and the
one-vertex fallback."""
import numpy as np

import oracle_py as O

POLY = np.array([[0.0, 0, 0], [0.25, 0, 0], [0.5, 0, 0], [0.75, 0, 0], [1.0, 0, 0]])


def test_foot_and_clamp():
    P = np.array([[0.3, 0.2, 0.1], [1.5, 0.1, 0.0], [-0.2, 0.0, 0.3]])
    o, a = O.line_projection(POLY, [[0, 1, 2, 3, 4]], [0], P, [0, 0, 0])
    assert np.allclose(o, [[0.3, 0, 0], [1.0, 0, 0], [0.0, 0, 0]], rtol=0, atol=1e-16)
    assert (a == [[1.0, 0, 0]] * 3).all()


def test_tie_goes_to_lower_segment():
    # polyline bent at (1,0,0): a point beyond the corner projects onto the corner from both
    # segments at the same distance; segment 0 (direction +x) must win
    V = np.array([[0.0, 0, 0], [1.0, 0, 0], [1.0, 1.0, 0]])
    P = np.array([[1.5, -0.5, 0.0], [1.2, 0.5, 0.0]])
    o, a = O.line_projection(V, [[0, 1, 2]], [0], P, [0, 0])
    assert (o[0] == [1.0, 0, 0]).all() and (a[0] == [1.0, 0, 0]).all()
    assert (o[1] == [1.0, 0.5, 0]).all() and (a[1] == [0.0, 1.0, 0]).all()
    # reversed vertex order: the +y segment is now first and wins the tie (direction -y)
    o, a = O.line_projection(V, [[2, 1, 0]], [0], P[:1], [0])
    assert (o[0] == [1.0, 0, 0]).all() and (a[0] == [0.0, -1.0, 0]).all()


def test_circle_closing_segment():
    V = np.array([[0.0, 0, 0], [1.0, 0, 0], [1.0, 1.0, 0], [0.0, 1.0, 0]])
    P = np.array([[-0.3, 0.4, 0.0]])
    o, a = O.line_projection(V, [[0, 1, 2, 3]], [1], P, [0])
    assert np.allclose(o[0], [0.0, 0.4, 0]) and (a[0] == [0.0, -1.0, 0]).all()
    # the same curve left open has no closing segment: nearest is vertex 0 or vertex 3
    o, a = O.line_projection(V, [[0, 1, 2, 3]], [0], P, [0])
    assert (o[0] == [0.0, 1.0, 0]).all() or (o[0] == [0.0, 0.0, 0]).all()


def test_short_segment_projects_to_first_vertex():
    V = np.array([[0.0, 0, 0], [1e-4, 0, 0]])          # |e|^2 = 1e-8 < 1e-7: t = 0
    o, a = O.line_projection(V, [[0, 1]], [0], np.array([[1.0, 1.0, 0.0]]), [0])
    assert (o[0] == [0.0, 0, 0]).all() and np.allclose(a[0], [1.0, 0, 0])


def test_single_vertex_fallback():
    V = np.array([[0.5, 0.5, 0.5]])
    o, a = O.line_projection(V, [[0]], [0], np.array([[0.0, 0.0, 0.0]]), [0])
    # nearest vertex; the tangent adds (0,0,0)/0 to (1,0,0), as the reference does
    assert (o[0] == V[0]).all() and np.isnan(a[0]).all()


def test_point_line_projection_vs_compiled_reference():
    """The oracle's point_line_projection against the reference's own, compiled from
    global_functions.cpp (tests/golden/point_line_cases.npz, tests/golden/make_lines.py):
    1,276 (segment, point) triples, including squared lengths around the 1e-7 threshold,
    zero-length segments, points at and perpendicular to the ends, and all of it shifted
    far from the origin. Bit for bit."""
    import os
    c = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "point_line_cases.npz"))
    pv = O.point_line_projection(c["A"], c["B"], c["P"])
    assert np.array_equal(pv, c["pv"])
    # both sides of the threshold and both clamps occur
    L2 = ((c["B"] - c["A"]) ** 2).sum(axis=1)
    assert (L2 < 1e-7).any() and ((L2 >= 1e-7) & (L2 < 1.1e-7)).any()
    assert (c["t"] < 0).any() and (c["t"] > 1).any() and (c["t"] == 0).any()
