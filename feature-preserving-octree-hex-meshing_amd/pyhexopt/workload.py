"""The clipped-torus feature-alignment workload on an adaptive octree hex mesh.

Input surface: the torus R = 1, r = 0.35 cut by the planes z = +-0.2, 40,000 triangles in
four patches (outer tube band, inner tube band, top and bottom annuli) meeting at four sharp
rim circles (meshgen.clipped_torus_surface).

Hex mesh (octree_mesh, grid_hex_meshing.cpp:460-567, then the inside test of clean_hex_mesh,
:1937-1952): the graded, paired octree of the surface (hexopt_octree_build) at num_voxels
voxels along the longest extent and STOP extent stop_extent, every leaf a hex; the leaves
whose centre has a negative pseudonormal signed distance (points_inside_mesh) are kept (the
cleaning passes are left out: they connect hexes through shared faces, which leaves of
different sizes do not have). Leaves of different sizes meet at 2:1
transitions, whose hanging nodes stay vertices of the smaller leaves only; the reference
turns such leaves into a conforming mesh by its dual construction (dual_octree_meshing,
:297-459), which is not built here.

Feature constraints, as feature_alignment / dirty_graph_projection set them up: boundary
vertices within 0.75 of their smallest incident leaf's edge length of a rim are LINE
vertices of that rim (curve projection); the other boundary vertices are REGULAR vertices
projected onto the patch of their nearest surface triangle (tri_Trees[patch]). The torus has
no corners, so no vertex is dirty.

Needs a GPU (octree build and closest points run on the device).
"""
import numpy as np

from . import capi, meshgen as mg

R_MAJ, R_MIN, Z_CUT = 1.0, 0.35, 0.2
SURFACE = dict(nu=200, nband=30, nring=20)


def torus_surface():
    """(V, F, rims, patch of every face): the faces come out band by band and annulus by
    annulus from clipped_torus_surface, so a face's patch follows from its index."""
    SV, SF, rims = mg.clipped_torus_surface(R_MAJ, R_MIN, Z_CUT, **SURFACE)
    nu, nb, nr = SURFACE["nu"], SURFACE["nband"], SURFACE["nring"]
    sizes = [2 * nu * nb, 2 * nu * nb, 2 * nu * nr, 2 * nu * nr]
    patch = np.repeat(np.arange(4), sizes).astype(np.int32)
    assert patch.size == SF.shape[0]
    return SV, SF, rims, patch


def clipped_torus_octree(num_voxels=1024, stop_extent=1, device=0):
    SV, SF, rims, fpatch = torus_surface()
    vs, origin, grid, mt = mg.octree_frame(SV, num_voxels)
    surf = capi.Surface(SV, SF, device=device)
    try:
        out = surf.octree(grid, origin, vs, mt, stop_extent)
        leaves, V, H = out["leaves"], out["V"], out["H"]
        cen = V[H].mean(axis=1)                   # leaves are boxes: the bbox centre
        S, _, _, _ = surf.closest(cen)
        inside = S < 0
        leaves, H = leaves[inside], H[inside]
        used = np.unique(H)
        remap = np.full(V.shape[0], -1, dtype=np.int64)
        remap[used] = np.arange(used.size)
        V, H = V[used], remap[H].astype(np.int32)
        quads = mg.octree_leaf_boundary(leaves, H)
        bnd = np.unique(quads)
        # smallest incident leaf edge length per vertex
        edge = np.full(V.shape[0], np.inf)
        for k in range(8):
            np.minimum.at(edge, H[:, k], leaves[:, 3] * vs)
        P = V[bnd]
        rho = np.sqrt(P[:, 0] ** 2 + P[:, 1] ** 2)
        w = np.sqrt(R_MIN ** 2 - Z_CUT ** 2)
        rim_geom = [(R_MAJ - w, -Z_CUT), (R_MAJ - w, Z_CUT), (R_MAJ + w, -Z_CUT), (R_MAJ + w, Z_CUT)]
        d = np.stack([np.hypot(rho - rr, P[:, 2] - zz) for rr, zz in rim_geom], axis=1)
        k = d.argmin(axis=1)
        line = d[np.arange(k.size), k] < 0.75 * edge[bnd]
        idsT = bnd[~line].astype(np.int32)
        _, I, _, _ = surf.closest(V[idsT])
        patchT = fpatch[I].astype(np.int32)
    finally:
        surf.close()
    return dict(V=V, H=H, leaves=leaves, quads=quads, idsT=idsT, patchT=patchT, idsL=bnd[line].astype(np.int32),
                curveL=k[line].astype(np.int32), SV=SV, SF=SF, fpatch=fpatch, curves=[list(r) for r in rims],
                voxel_size=vs, grid=grid, octree_leaves=int(out["leaves"].shape[0]), inside=int(inside.sum()),
                extents=np.unique(leaves[:, 3], return_counts=True))


def patch_surfaces(w, device=0):
    """One Surface per patch (its triangles of the input surface), in patch order."""
    out = []
    for p in range(4):
        F = w["SF"][w["fpatch"] == p]
        used = np.unique(F)
        remap = np.full(w["SV"].shape[0], -1, dtype=np.int64)
        remap[used] = np.arange(used.size)
        out.append(capi.Surface(w["SV"][used], remap[F].astype(np.int32), device=device))
    return out


def box_union_outer(V, F, boxes, nper):
    """The triangles of the box-union soup on the union's outer surface: those whose centroid
    lies in no other box (a box's own faces have their centroids on its boundary)."""
    cen = V[F].mean(axis=1)
    owner = np.arange(F.shape[0]) // nper
    keep = np.ones(F.shape[0], dtype=bool)
    for b, (Rm, sides, centre) in enumerate(boxes):
        q = np.abs((cen - centre) @ Rm) <= sides / 2 + 1e-12
        keep &= ~(q.all(axis=1) & (owner != b))
    return F[keep]


def box_union_octree(num_voxels=512, stop_extent=1, device=0, nboxes=8, n=46, seed=0):
    """The scale workload: the union of `nboxes` seeded rotated boxes (meshgen.box_union_surface,
    ~203k triangles as the soup of the closed box surfaces), its graded, paired octree, and the
    leaves inside the union by the signed +z crossing count of compute_sign (voxelization.cpp:
    101-163, hexopt_occ_cell_parity), which makes the soup the union. Boundary vertices are
    REGULAR vertices projected onto the union's outer triangles (box_union_outer); there are
    no feature curves in this workload."""
    SV, SF, boxes = mg.box_union_surface(nboxes, n, seed)
    nper = SF.shape[0] // nboxes
    vs, origin, grid, mt = mg.octree_frame(SV, num_voxels)
    surf = capi.Surface(SV, SF, device=device)
    try:
        out = surf.octree(grid, origin, vs, mt, stop_extent)
        leaves, V, H = out["leaves"], out["V"], out["H"]
        # compute_sign's octree variant on the leaves, in the frame the octree was built in
        inside = surf.cell_parity(leaves, origin + mt, vs) > 0.5
    finally:
        surf.close()
    leaves, H = leaves[inside], H[inside]
    used = np.unique(H)
    remap = np.full(V.shape[0], -1, dtype=np.int64)
    remap[used] = np.arange(used.size)
    V, H = V[used], remap[H].astype(np.int32)
    quads = mg.octree_leaf_boundary(leaves, H)
    bnd = np.unique(quads).astype(np.int32)
    SFo = box_union_outer(SV, SF, boxes, nper)
    return dict(V=V, H=H, leaves=leaves, quads=quads, idsT=bnd, patchT=np.zeros(bnd.size, np.int32),
                idsL=np.zeros(0, np.int32), curveL=np.zeros(0, np.int32), SV=SV, SF=SFo, SF_soup=SF,
                fpatch=np.zeros(SFo.shape[0], np.int32), curves=[], voxel_size=vs, grid=grid,
                octree_leaves=int(out["leaves"].shape[0]), inside=int(inside.sum()),
                extents=np.unique(leaves[:, 3], return_counts=True))


def sampled_hausdorff(V0, F0, V1, F1, n=1_000_000, seed=0, device=0):
    """Hausdorff distance by area-uniform sampling: n seeded points on each triangle mesh
    (meshgen.surface_samples), each measured against the other mesh by the device closest
    point (hexopt_surface_closest). Returns dict(max, mean) per direction ("0to1", "1to0")
    and the symmetric maximum."""
    out = {}
    for tag, (Va, Fa, Vb, Fb, sd) in (("0to1", (V0, F0, V1, F1, seed)), ("1to0", (V1, F1, V0, F0, seed + 1))):
        P = mg.surface_samples(Va, Fa, n, sd)
        s = capi.Surface(Vb, Fb, device=device)
        try:
            S, _, _, _ = s.closest(P)
        finally:
            s.close()
        d = np.abs(S)
        out[tag] = dict(max=float(d.max()), mean=float(d.mean()))
    out["max"] = max(out["0to1"]["max"], out["1to0"]["max"])
    return out
