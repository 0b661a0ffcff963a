"""The element passes on meshes other than consistently numbered grids: a jittered 14^3 grid
with half its hexes numbered from rotated corners (most tiles scatter by hex colour) and
the graded octree leaf mesh of the clipped torus (hanging nodes at the 2:1 transitions).
dE/dx (FP64 and FP32) against the oracle, the global operator L p against the oracle's, and
the FP32 optimiser's FP64 residual against its own CG estimate."""
import numpy as np
import pytest

import oracle_py as O
from helpers import rel

pytestmark = pytest.mark.gpu


def _rotated():
    from pyhexopt import meshgen as mg
    n = 14
    V, H = mg.grid_hex(n)
    V = mg.jitter(V, ~mg.grid_boundary_mask(n), 0.15 * 2.0 / n, seed=8)
    return V, mg.rotate_hex_corners(H, 0.5, seed=8)


def _octree():
    from pyhexopt import workload
    w = workload.clipped_torus_octree(64, 1)
    V = w["V"] + 0.1 * w["voxel_size"] * np.random.default_rng(3).normal(size=w["V"].shape)
    return V, w["H"]


@pytest.fixture(params=["rotated", "octree"])
def mesh(request):
    V, H = _rotated() if request.param == "rotated" else _octree()
    print("%s mesh: %d hexes, %d vertices" % (request.param, H.shape[0], V.shape[0]), flush=True)
    return V, H


def test_gradients_and_operator_vs_oracle(mesh):
    from pyhexopt import capi
    V, H = mesh
    h = capi.HexOpt(V, H)
    try:
        h.reference()
        info = h.layout_info()
        X = V + 1e-3 * np.random.default_rng(1).normal(size=V.shape) * np.abs(V).max() / 50
        g64 = h.energy_grad(X)
        g32 = h.energy_grad(X.astype(np.float32), precision=32).astype(np.float64)
        g64f = h.energy_grad(X.astype(np.float32).astype(np.float64))
        h.set_positions(capi.CURRENT, X)
        p = np.random.default_rng(2).normal(size=3 * V.shape[0])
        _, _, Lp = h.debug_system(p)
    finally:
        h.close()
    G = O.grad_ref(V, H)
    _, go = O.sd_grad(H, G, X)
    _, _, Lpo, _ = O.linear_system(V, H, G, X, dense=False, p=p)
    print("tiles %d (corner %d): grad64 %.2e, grad32 %.2e, L p %.2e" %
          (info["tiles"], info["corner_tiles"], rel(g64, go), rel(g32, g64f), rel(Lp, Lpo)), flush=True)
    assert rel(g64, go) < 1e-10
    # FP32 where I + D is far from isometry (the jittered grid's worst tets) keeps ~1e-5
    assert rel(g32, g64f) < 5e-5
    assert rel(Lp, Lpo) < 1e-10


def test_fp32_residual_is_the_cg_estimate(mesh):
    from pyhexopt import capi
    V, H = mesh
    h = capi.HexOpt(V, H)
    try:
        h.reference()
        _, log = h.opt32_solve(1, residual_every=1)
    finally:
        h.close()
    print("FP32 CG estimate %.4e, FP64 residual %.4e" % (log[0, 6], log[0, 7]), flush=True)
    assert abs(log[0, 7] - log[0, 6]) <= 0.05 * log[0, 6] + 1e-6


@pytest.mark.parametrize("kinds", ["planes", "lines", "planes+lines"])
def test_fp32_residual_with_feature_constraints(kinds):
    """The feature-alignment set-up of the octree torus (patch planes, rim lines projected on
    the device): the FP32 optimiser's FP64 residual against its CG estimate."""
    from pyhexopt import capi, workload
    w = workload.clipped_torus_octree(64, 1)
    h = capi.HexOpt(w["V"], w["H"])
    patches = workload.patch_surfaces(w)
    curves = capi.Curves(w["SV"], w["curves"], np.ones(len(w["curves"]), dtype=np.uint8))
    try:
        kw = {}
        if "planes" in kinds:
            nT = w["idsT"].size
            kw.update(idsT=w["idsT"], normalT=np.zeros((nT, 3)), disT=np.zeros(nT), lamT=0.05)
        if "lines" in kinds:
            nL = w["idsL"].size
            kw.update(idsL=w["idsL"], axaL=np.zeros((nL, 3)), originL=np.zeros((nL, 3)), lamL=0.05)
        h.set_constraints(**kw)
        h.reference()
        if "planes" in kinds:
            h.project_patches(patches, w["patchT"], 0.05)
        if "lines" in kinds:
            h.refresh_lines(curves, w["curveL"], 0.05)
        _, log = h.opt32_solve(1, weight_opt=0.001, residual_every=1)
    finally:
        h.close()
        for p in patches:
            p.close()
        curves.close()
    print("%s: FP32 CG estimate %.4e, FP64 residual %.4e" % (kinds, log[0, 6], log[0, 7]), flush=True)
    assert abs(log[0, 7] - log[0, 6]) <= 0.05 * log[0, 6] + 1e-6
