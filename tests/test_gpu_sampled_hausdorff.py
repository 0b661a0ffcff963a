"""Hausdorff distance over area-uniform surface samples (workload.sampled_hausdorff, the
scale workload's 1M-sample measure): the device closest-point distances of the samples
against brute force over every triangle (the oracle's exact point-triangle distance) on a
small case."""
import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu


def test_sampled_distances_vs_brute_force():
    from pyhexopt import capi, meshgen as mg, workload
    SV, SF = mg.superellipsoid(2)
    V, H = mg.grid_hex(4, -0.8, 0.8)
    T = mg.hex_boundary_triangles(H)
    res = workload.sampled_hausdorff(SV, SF, V, T, n=3000, seed=4)
    for tag, (Va, Fa, Vb, Fb, sd) in (("0to1", (SV, SF, V, T, 4)), ("1to0", (V, T, SV, SF, 5))):
        P = mg.surface_samples(Va, Fa, 3000, sd)
        FN, VN, FEN = capi.surface_normals(Vb, Fb)
        S, _, _, _, _ = O.signed_distance_pseudonormal(P, Vb, Fb, FN, VN, FEN)
        d = np.abs(S)
        print("%s: max %.6e (brute force %.6e)" % (tag, res[tag]["max"], d.max()), flush=True)
        assert res[tag]["max"] == pytest.approx(d.max(), rel=1e-12)
        assert res[tag]["mean"] == pytest.approx(d.mean(), rel=1e-12)
