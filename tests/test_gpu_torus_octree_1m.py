"""The FP32+adjoint optimiser against the FP64 plain path at full size: the 1,135,816-hex
graded octree mesh of the clipped torus (num_voxels 448), one iteration of the
feature-alignment loop (grid_hex_meshing.cpp:3855-3869: compute_referenceMesh, the patch
tangent planes of dirty_graph_projection :4050-4080, the rim lines, one SLIM iteration,
lamda 0.05, weight_opt 0.001) from the same start.

Bounds, from the errors the FP32 arithmetic leaves (measured values in brackets):
- dE/dx at the same (FP32-representable) positions: 1e-6 of the largest component [3.6e-8];
- SLIM weight blocks: 1e-6 of the largest entry [2.8e-8];
- the FP32 CG (the reference's 20 steps, no refinement) reproduces the FP64 path's truncated
  solve: directions within 1e-5 relative [1.1e-7], the same accepted step of the flip-free
  line search (the first iteration on this mesh is the case whose FP32-summed line-search
  energies once rejected the FP64 path's step), quality energy after the step within 1e-12
  [8e-16];
- the FP64 residual of the FP32 direction agrees with the FP32 CG's own estimate to 1e-6
  [1.6e-8]."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WEIGHT_OPT, LAM = 0.001, 0.05


@pytest.fixture(scope="module")
def torus():
    from pyhexopt import workload
    w = workload.clipped_torus_octree(448, 1)
    print("octree torus: %d hexes, %d vertices, %d plane + %d line vertices" %
          (w["H"].shape[0], w["V"].shape[0], w["idsT"].size, w["idsL"].size), flush=True)
    return w


def _context(w):
    from pyhexopt import capi, workload
    h = capi.HexOpt(w["V"], w["H"])
    nT, nL = w["idsT"].size, w["idsL"].size
    h.set_constraints(idsT=w["idsT"], normalT=np.zeros((nT, 3)), disT=np.zeros(nT), lamT=LAM,
                      idsL=w["idsL"], axaL=np.zeros((nL, 3)), originL=np.zeros((nL, 3)), lamL=LAM)
    patches = workload.patch_surfaces(w)
    curves = capi.Curves(w["SV"], w["curves"], np.ones(len(w["curves"]), dtype=np.uint8))
    h.rest_from_current()
    h.reference()
    h.project_patches(patches, w["patchT"], LAM)
    h.refresh_lines(curves, w["curveL"], LAM)
    return h, patches, curves


def _close(ctx):
    h, patches, curves = ctx
    h.close()
    for p in patches:
        p.close()
    curves.close()


def test_first_iteration_fp32_vs_fp64(torus):
    w = torus
    c64, c32 = _context(w), _context(w)
    try:
        h64, h32 = c64[0], c32[0]
        X0 = h64.get_positions()
        # both gradients at the same (FP32-representable) positions: the FP32 adjoint's own error
        X32 = X0.astype(np.float32)
        g64 = h64.energy_grad(X32.astype(np.float64), precision=64, weight_opt=WEIGHT_OPT)
        g32 = h64.energy_grad(X32, precision=32, weight_opt=WEIGHT_OPT).astype(np.float64)
        eg = np.abs(g32 - g64).max() / np.abs(g64).max()
        K64 = h64.debug_operator(64, weight_opt=WEIGHT_OPT)
        K32 = h64.debug_operator(32, weight_opt=WEIGHT_OPT)
        ek = np.abs(K32 - K64).max() / np.abs(K64).max()
        print("dE/dx %.2e, weight blocks %.2e" % (eg, ek), flush=True)
        _, _, _, log64, _ = h64.slim_solve(1, weight_opt=WEIGHT_OPT)
        _, log32 = h32.opt32_solve(1, cg_iters=20, weight_opt=WEIGHT_OPT, refine_iters=0, residual_every=1)
        s64, s32 = float(log64[0, 4]), float(log32[0, 4])
        d64 = (h64.get_positions() - X0) / s64
        d32 = (h32.get_positions() - X0) / s32
        ed = np.linalg.norm(d32 - d64) / np.linalg.norm(d64)
        ee = abs(log32[0, 1] - log64[0, 1]) / abs(log64[0, 1])
        er = abs(log32[0, 7] - log32[0, 6]) / log32[0, 6]
        print("steps %g / %g, direction %.2e, quality energy %.2e, FP64 residual %.4e vs CG estimate %.4e" %
              (s32, s64, ed, ee, log32[0, 7], log32[0, 6]), flush=True)
    finally:
        _close(c64)
        _close(c32)
    assert eg < 1e-6 and ek < 1e-6
    assert s32 == s64 and s64 > 0
    assert ed < 1e-5
    assert ee < 1e-12
    assert er < 1e-6
