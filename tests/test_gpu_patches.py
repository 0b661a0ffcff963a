"""hexopt_project_patches (dirty_graph_projection's per-patch projection,
grid_hex_meshing.cpp:4050-4080) on the graded octree mesh of the clipped torus, whose four
patches are projected in one launch: the system and two SLIM iterations are bit for bit those
that hexopt_set_constraints gives with the tangent planes formed from each patch's own
closest-point queries (hexopt_surface_closest on that patch's surface), at the lattice start
and at jittered positions."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("jitter", [0.0, 0.1])
def test_project_patches_equals_per_patch_planes(jitter):
    from pyhexopt import capi, workload
    w = workload.clipped_torus_octree(64, 1)
    V = w["V"] + jitter * w["voxel_size"] * np.random.default_rng(5).normal(size=w["V"].shape)
    ids, pT = w["idsT"], w["patchT"]
    lam = 0.05
    patches = workload.patch_surfaces(w)
    h1, h2 = capi.HexOpt(V, w["H"]), capi.HexOpt(V, w["H"])
    try:
        print("octree torus: %d hexes, %d plane vertices over %d patches, jitter %g" %
              (w["H"].shape[0], ids.size, len(patches), jitter), flush=True)
        for h in (h1, h2):
            h.reference()
        nT = ids.size
        h1.set_constraints(idsT=ids, normalT=np.zeros((nT, 3)), disT=np.zeros(nT), lamT=1.0)
        h1.project_patches(patches, pT, lam)
        # a second call reusing the assignment gives the same planes
        h1.project_patches(patches, None, lam)
        N = np.zeros((nT, 3))
        Cp = np.zeros((nT, 3))
        for p, s in enumerate(patches):
            sel = np.nonzero(pT == p)[0]
            _, _, C, Np = s.closest(V[ids[sel]])
            N[sel], Cp[sel] = Np, C
        dis = N[:, 0] * Cp[:, 0] + N[:, 1] * Cp[:, 1] + N[:, 2] * Cp[:, 2]
        h2.set_constraints(idsT=ids, normalT=N, disT=dis, lamT=lam)
        x = np.random.default_rng(0).normal(size=3 * V.shape[0])
        for a, b in zip(h1.debug_system(x, weight_opt=0.001), h2.debug_system(x, weight_opt=0.001)):
            assert np.array_equal(a, b)
        assert h1.energy(weight_opt=0.001) == h2.energy(weight_opt=0.001)
        r1, r2 = h1.slim_solve(2, weight_opt=0.001), h2.slim_solve(2, weight_opt=0.001)
        # without a previous assignment (after new constraints) None is refused
        h3 = capi.HexOpt(V, w["H"])
        try:
            h3.set_constraints(idsT=ids, normalT=np.zeros((nT, 3)), disT=np.zeros(nT), lamT=1.0)
            with pytest.raises(RuntimeError):
                h3.project_patches(patches, None, lam)
        finally:
            h3.close()
        assert r1[0] == r2[0]
        assert np.array_equal(h1.get_positions(), h2.get_positions())
    finally:
        h1.close()
        h2.close()
        for s in patches:
            s.close()
