"""Writes tests/golden/point_line_cases.npz: the reference's point_line_projection
(global_functions.cpp:3454-3465, compiled from the reference's sources into
oracle/_ref/slim_ref, driver mode 6) on (segment, point) triples: random segments and
points on both sides of the ends, segments around the Precision = 1e-7 threshold on the
squared length, zero-length segments, points at and perpendicular to the ends, and the same
geometry shifted far from the origin. Run in the development container:
  python tests/golden/make_lines.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import refio  # noqa: E402


def cases():
    rng = np.random.default_rng(11)
    A, B, P = [], [], []
    # random segments, points with t in about [-0.5, 1.5] and off the line
    a = rng.normal(size=(400, 3)); b = a + rng.normal(size=(400, 3))
    t = rng.uniform(-0.5, 1.5, size=(400, 1))
    A.append(a); B.append(b); P.append(a + t * (b - a) + 0.3 * rng.normal(size=(400, 3)))
    # squared length around the 1e-7 threshold
    L = np.sqrt(1e-7) * np.exp(rng.uniform(-0.01, 0.01, size=200))
    d = rng.normal(size=(200, 3)); d /= np.linalg.norm(d, axis=1)[:, None]
    a = rng.normal(size=(200, 3))
    A.append(a); B.append(a + L[:, None] * d); P.append(a + 1e-3 * rng.normal(size=(200, 3)))
    # zero-length segments
    a = rng.normal(size=(20, 3))
    A.append(a); B.append(a.copy()); P.append(a + rng.normal(size=(20, 3)))
    # points at the ends and perpendicular to them
    a = np.array([[0.0, 0, 0], [1.0, 2, 3], [-1.0, 0.5, 0.25]]); b = a + np.array([[1.0, 0, 0], [0, 0.5, 0], [0.25, 0.25, 0]])
    n = np.array([[0.0, 1, 0], [1.0, 0, 0], [0.0, 0, 1.0]])
    for pts in (a, b, a + n, b + n, a - (b - a), b + (b - a)):
        A.append(a); B.append(b); P.append(pts)
    A0, B0, P0 = (np.concatenate(x) for x in (A, B, P))
    # the same far from the origin
    off = np.array([1e3, -2e3, 5e2])
    return np.concatenate([A0, A0 + off]), np.concatenate([B0, B0 + off]), np.concatenate([P0, P0 + off])


def main():
    if not refio.available():
        sys.exit("oracle/_ref/slim_ref is not built")
    A, B, P = cases()
    o = refio.run_point_line(A, B, P)
    np.savez(os.path.join(HERE, "point_line_cases.npz"), A=A, B=B, P=P, pv=o["pv"].reshape(-1, 3), t=o["t"])
    print("%d triples" % P.shape[0])


if __name__ == "__main__":
    main()
