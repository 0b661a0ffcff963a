"""Synthetic workloads for the hex-optimisation path.

Hex corner order is the reference's `Hybrid::vs` order (global_types.h:216-226,
`hex_ref_shape`): 0=(0,0,0) 1=(1,0,0) 2=(1,1,0) 3=(0,1,0) 4=(0,0,1) 5=(1,0,1) 6=(1,1,1)
7=(0,1,1). Vertex arrays are (nV, 3) float64 row-major unless stated otherwise.
"""
import numpy as np

HEX_CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                        [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.int64)
# global_types.h:163-173
HEX_TETRA_TABLE = np.array([[0, 3, 4, 1], [1, 0, 5, 2], [2, 1, 6, 3], [3, 2, 7, 0],
                            [4, 7, 5, 0], [5, 4, 6, 1], [6, 5, 7, 2], [7, 6, 4, 3]], dtype=np.int64)


def grid_hex(n, lo=-1.0, hi=1.0):
    """Regular n^3 hex grid on [lo,hi]^3. Returns V (nV,3), H (nH,8) int32."""
    t = np.linspace(lo, hi, n + 1)
    z, y, x = np.meshgrid(t, t, t, indexing="ij")
    V = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.float64)
    idx = np.arange((n + 1) ** 3).reshape(n + 1, n + 1, n + 1)  # [z, y, x]
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    H = np.stack([idx[k + c[2], j + c[1], i + c[0]].ravel() for c in HEX_CORNERS], axis=1)
    return V, H.astype(np.int32)


def grid_boundary_mask(n):
    t = np.arange(n + 1)
    z, y, x = np.meshgrid(t, t, t, indexing="ij")
    b = (x == 0) | (x == n) | (y == 0) | (y == n) | (z == 0) | (z == n)
    return b.ravel()


def jitter(V, interior_mask, sigma, seed=0):
    rng = np.random.default_rng(seed)
    out = V.copy()
    ids = np.nonzero(interior_mask)[0]
    out[ids] += rng.normal(0.0, sigma, size=(ids.size, 3))
    return out


def invert_hexes(V, H, frac, seed=0):
    """Swap the positions of corners 0 and 1 in a seeded fraction of hexes (inverts them).
    Hexes sharing a vertex may be picked; the swap is applied in hex order."""
    rng = np.random.default_rng(seed)
    nH = H.shape[0]
    k = int(round(frac * nH))
    pick = rng.choice(nH, size=k, replace=False)
    out = V.copy()
    for h in np.sort(pick):
        a, b = H[h, 0], H[h, 1]
        out[[a, b]] = out[[b, a]]
    return out, np.sort(pick)


def cube_rotations():
    """The 24 proper rotations of the cube as corner permutations: hex H[:, perm] is the same
    hex, positively oriented, with its corners numbered from another corner and axis."""
    C = 2 * HEX_CORNERS - 1
    key = {tuple(c): k for k, c in enumerate(C.tolist())}
    out = []
    for p in ([0, 1, 2], [1, 2, 0], [2, 0, 1], [0, 2, 1], [2, 1, 0], [1, 0, 2]):
        for sx in (1, -1):
            for sy in (1, -1):
                for sz in (1, -1):
                    R = np.zeros((3, 3), dtype=np.int64)
                    R[0, p[0]], R[1, p[1]], R[2, p[2]] = sx, sy, sz
                    if round(np.linalg.det(R)) != 1:
                        continue
                    out.append([key[tuple((R @ c).tolist())] for c in C])
    return np.array(out, dtype=np.int64)


def rotate_hex_corners(H, frac, seed=0):
    """Renumber the corners of a seeded fraction of hexes by random proper cube rotations
    (other than the identity). The mesh, each hex's orientation and its 8 corner tets are
    unchanged; only which corner is called 0 differs."""
    rng = np.random.default_rng(seed)
    rots = cube_rotations()[1:]
    out = H.copy()
    pick = rng.choice(H.shape[0], size=int(round(frac * H.shape[0])), replace=False)
    which = rng.integers(0, rots.shape[0], size=pick.size)
    out[pick] = np.take_along_axis(H[pick], rots[which], axis=1)
    return out


def exploded_start(V, H, fixed, target, amax, seed=0):
    """Start positions of a padding-untangle glue (one copy of each corner per hex, (nH, 8, 3)),
    in the manner of optimization::rigidDeformCubes (optimization.cpp:457-497): a hex with a
    pinned vertex starts at its target corners; any other hex starts as the hex's shape in V,
    turned by a seeded random rotation of up to amax radians and scaled by up to 20% about its
    centre, placed at the centre of its target corners."""
    rng = np.random.default_rng(seed)
    out = np.asarray(V, dtype=np.float64)[H].copy()
    T = np.asarray(target, dtype=np.float64)
    hb = np.asarray(fixed, dtype=bool)[H].any(axis=1)
    for h in range(H.shape[0]):
        if hb[h]:
            out[h] = T[H[h]]
            continue
        ax = rng.normal(size=3)
        ax /= np.linalg.norm(ax)
        K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
        a = rng.uniform(-amax, amax)
        R = np.eye(3) + np.sin(a) * K + (1 - np.cos(a)) * K @ K
        c = out[h].mean(0)
        out[h] = (out[h] - c) @ R.T * rng.uniform(0.8, 1.2) + T[H[h]].mean(0)
    return out


def exploded_mesh(H, start):
    """Exploded mesh of hexes H with per-hex corner copies `start` (nH, 8, 3): vertices
    8h + i, hexes [8h .. 8h+7], and orig_of[8h + i] = H[h, i]."""
    nH = H.shape[0]
    Ve = np.asarray(start, dtype=np.float64).reshape(nH * 8, 3)
    He = np.arange(nH * 8, dtype=np.int32).reshape(nH, 8)
    return Ve, He, np.asarray(H, dtype=np.int32).reshape(-1).copy()


def icosphere(level):
    """Unit icosphere with 20*4^level triangles (outward orientation)."""
    p = (1.0 + 5 ** 0.5) / 2.0
    V = [[-1, p, 0], [1, p, 0], [-1, -p, 0], [1, -p, 0], [0, -1, p], [0, 1, p], [0, -1, -p], [0, 1, -p],
         [p, 0, -1], [p, 0, 1], [-p, 0, -1], [-p, 0, 1]]
    F = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
         [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
         [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    V = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in V]
    for _ in range(level):
        cache = {}
        nF = []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = V[a] + V[b]
                V.append(m / np.linalg.norm(m))
                cache[key] = len(V) - 1
            return cache[key]

        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nF += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        F = nF
    return np.array(V), np.array(F, dtype=np.int32)


def superellipsoid(level=5, power=4.0):
    """|x|^p+|y|^p+|z|^p = 1 surface: icosphere vertices mapped radially p/||p||_power."""
    V, F = icosphere(level)
    nrm = (np.abs(V) ** power).sum(axis=1) ** (1.0 / power)
    return V / nrm[:, None], F


def cube_surface(n=16, lo=0.0, hi=1.0):
    """Unit-cube surface, each face an n x n quad grid split along one diagonal
    (12*n^2 triangles, outward orientation)."""
    t = np.linspace(lo, hi, n + 1)
    verts = {}
    V = []

    def vid(p):
        key = tuple(np.round(p, 12))
        if key not in verts:
            verts[key] = len(V)
            V.append(p)
        return verts[key]

    F = []
    for axis in range(3):
        for side in (0, 1):
            u_ax, v_ax = [a for a in range(3) if a != axis]
            for i in range(n):
                for j in range(n):
                    q = []
                    for (di, dj) in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        p = np.zeros(3)
                        p[axis] = hi if side else lo
                        p[u_ax] = t[i + di]
                        p[v_ax] = t[j + dj]
                        q.append(vid(p))
                    a, b, c, d = q
                    tri = [[a, b, c], [a, c, d]]
                    # orient outward: normal of (a,b,c) along +axis when side==1
                    e1 = np.array(V[b]) - np.array(V[a])
                    e2 = np.array(V[c]) - np.array(V[a])
                    nz = np.cross(e1, e2)[axis]
                    if (nz > 0) != bool(side):
                        tri = [[a, c, b], [a, d, c]]
                    F += tri
    return np.array(V), np.array(F, dtype=np.int32)


def torus_surface(R=1.0, r=0.35, nu=200, nv=100):
    """Torus of radii (R, r) as an nu x nv parameter grid split into 2*nu*nv triangles."""
    u = np.arange(nu) * 2 * np.pi / nu
    v = np.arange(nv) * 2 * np.pi / nv
    uu, vv = np.meshgrid(u, v, indexing="ij")
    X = (R + r * np.cos(vv)) * np.cos(uu)
    Y = (R + r * np.cos(vv)) * np.sin(uu)
    Z = r * np.sin(vv)
    V = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    idx = np.arange(nu * nv).reshape(nu, nv)
    F = []
    for i in range(nu):
        for j in range(nv):
            a, b = idx[i, j], idx[(i + 1) % nu, j]
            c, d = idx[(i + 1) % nu, (j + 1) % nv], idx[i, (j + 1) % nv]
            F += [[a, b, c], [a, c, d]]
    return V, np.array(F, dtype=np.int32)


def voxel_hex_mesh(inside_fn, lo, hi, n):
    """Uniform-octree (single level) hex mesh: the cells of an n^3 grid on [lo,hi]^3 whose
    centre satisfies inside_fn. Unused vertices are dropped. Returns V, H."""
    V, H = grid_hex(n, lo, hi)
    cen = V[H].mean(axis=1)
    keep = inside_fn(cen)
    H = H[keep]
    used = np.unique(H)
    remap = -np.ones(V.shape[0], dtype=np.int64)
    remap[used] = np.arange(used.size)
    return V[used], remap[H].astype(np.int32)


def boundary_vertices(H, nV):
    """Vertices on boundary quads (quads used by exactly one hex)."""
    faces = np.array([[0, 1, 2, 3], [4, 7, 6, 5], [0, 4, 5, 1], [0, 3, 7, 4], [3, 2, 6, 7], [1, 5, 6, 2]])
    Q = H[:, faces].reshape(-1, 4)
    key = np.sort(Q, axis=1)
    _, inv, cnt = np.unique(key, axis=0, return_inverse=True, return_counts=True)
    bq = Q[cnt[inv.ravel()] == 1]
    mask = np.zeros(nV, dtype=bool)
    mask[bq.ravel()] = True
    return mask


def clipped_torus_inside(P, R=1.0, r=0.35, zcut=0.2):
    rho = np.sqrt(P[:, 0] ** 2 + P[:, 1] ** 2)
    return ((rho - R) ** 2 + P[:, 2] ** 2 <= r * r) & (np.abs(P[:, 2]) <= zcut)


def clipped_torus_closest(P, R=1.0, r=0.35, zcut=0.2):
    """Closest point and outward unit normal on the surface of the torus (R, r) cut by the
    planes z = +-zcut: the tube band, the two flat annuli and their four rim circles."""
    rho = np.maximum(np.sqrt(P[:, 0] ** 2 + P[:, 1] ** 2), 1e-300)
    ex, ey = P[:, 0] / rho, P[:, 1] / rho
    q = np.stack([rho - R, P[:, 2]], axis=1)
    qn = np.maximum(np.linalg.norm(q, axis=1), 1e-300)
    cands = []
    # tube band
    tq = q / qn[:, None] * r
    ok = np.abs(tq[:, 1]) <= zcut
    cands.append((tq, np.where(ok, np.abs(qn - r), np.inf), q / qn[:, None]))
    # annuli z = +-zcut, radial extent |rho - R| <= sqrt(r^2 - zcut^2)
    w = np.sqrt(r * r - zcut * zcut)
    for s in (1.0, -1.0):
        rr = np.clip(q[:, 0], -w, w)
        c = np.stack([rr, np.full_like(rr, s * zcut)], axis=1)
        d = np.linalg.norm(q - c, axis=1)
        n = np.tile([0.0, s], (q.shape[0], 1))
        inside_ann = np.abs(q[:, 0]) <= w
        d = np.where(inside_ann, np.abs(q[:, 1] - s * zcut), d)
        # off the annulus the nearest point is a rim circle; its normal is the direction to p
        rim = ~inside_ann
        dn = q - c
        dn = dn / np.maximum(np.linalg.norm(dn, axis=1), 1e-300)[:, None]
        n = np.where(rim[:, None], dn, n)
        cands.append((c, d, n))
    best = np.argmin(np.stack([cd[1] for cd in cands], axis=1), axis=1)
    C2 = np.zeros_like(q)
    N2 = np.zeros_like(q)
    for k, (c, d, n) in enumerate(cands):
        m = best == k
        C2[m] = c[m]
        N2[m] = n[m]
    C = np.stack([(R + C2[:, 0]) * ex, (R + C2[:, 0]) * ey, C2[:, 1]], axis=1)
    N = np.stack([N2[:, 0] * ex, N2[:, 0] * ey, N2[:, 1]], axis=1)
    N /= np.maximum(np.linalg.norm(N, axis=1), 1e-300)[:, None]
    return C, N


def torus_hex_mesh(target_hexes=1_000_000, R=1.0, r=0.35, zcut=0.2):
    """Uniform-level octree hex mesh of the clipped torus: cells of a cubic grid whose centre
    lies inside (cell size chosen for ~target_hexes). Returns V, H, boundary vertex ids."""
    vol = 2 * np.pi * R * (np.pi * r * r) * 0.684   # clipped-torus volume fraction, calibrated
    h = (vol / target_hexes) ** (1.0 / 3.0)
    ext = R + r + h
    n = int(np.ceil(2 * ext / h))
    V, H = voxel_hex_mesh(lambda P: clipped_torus_inside(P, R, r, zcut), -ext, -ext + n * h, n)
    bnd = np.nonzero(boundary_vertices(H, V.shape[0]))[0].astype(np.int32)
    return V, H, bnd


def clipped_torus_surface(R=1.0, r=0.35, zcut=0.2, nu=200, nband=20, nring=10):
    """Watertight triangulation of the torus (R, r) cut by the planes z = +-zcut: the outer
    and inner tube bands (nu x nband quads each), the two flat annuli (nu x nring quads
    each) and the four shared rim circles (nu vertices each). Outward orientation.
    Returns V, F and the rim vertex lists (inner bottom, inner top, outer bottom, outer top)."""
    a = np.arcsin(zcut / r)
    w = np.sqrt(r * r - zcut * zcut)
    u = np.arange(nu) * 2 * np.pi / nu
    V = []

    def ring(rad, z):
        base = len(V)
        V.extend([[rad * np.cos(t), rad * np.sin(t), z] for t in u])
        return np.arange(base, base + nu)

    # tube angle v: outer band v in [-a, a], inner band v in [pi - a, pi + a]
    def band(v0, v1, first, last):
        rows = [first]
        for k in range(1, nband):
            v = v0 + (v1 - v0) * k / nband
            rows.append(ring(R + r * np.cos(v), r * np.sin(v)))
        rows.append(last)
        return rows

    ib, it = ring(R - w, -zcut), ring(R - w, zcut)
    ob, ot = ring(R + w, -zcut), ring(R + w, zcut)
    F = []

    def strip(rows):
        for lo, hi in zip(rows[:-1], rows[1:]):
            for i in range(nu):
                j = (i + 1) % nu
                F.append([lo[i], lo[j], hi[j]])
                F.append([lo[i], hi[j], hi[i]])

    strip(band(-a, a, ob, ot))                      # outer band, bottom rim to top rim
    strip(band(np.pi - a, np.pi + a, it, ib))       # inner band, top rim to bottom rim
    # annuli: rings from the inner rim to the outer rim
    for z, first, last, up in ((zcut, it, ot, True), (-zcut, ib, ob, False)):
        rows = [first] + [ring(R - w + 2 * w * k / nring, z) for k in range(1, nring)] + [last]
        m0 = len(F)
        strip(rows)
        if not up:
            for f in F[m0:]:
                f[1], f[2] = f[2], f[1]
    V = np.array(V)
    F = np.array(F, dtype=np.int32)
    # orient every face outward: band faces by the tube normal, annuli by +-z
    C = V[F].mean(axis=1)
    n = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    rho = np.sqrt(C[:, 0] ** 2 + C[:, 1] ** 2)
    out = np.stack([C[:, 0] / rho * (rho - R), C[:, 1] / rho * (rho - R), C[:, 2]], axis=1)
    flat = np.abs(np.abs(C[:, 2]) - zcut) < 1e-12
    out[flat] = np.stack([np.zeros(flat.sum()), np.zeros(flat.sum()), np.sign(C[flat, 2])], axis=1)
    flip = (n * out).sum(axis=1) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return V, F, [ib, it, ob, ot]


HEX_FACE_TABLE = np.array([[0, 1, 2, 3], [4, 7, 6, 5], [0, 4, 5, 1], [0, 3, 7, 4], [3, 2, 6, 7], [1, 5, 6, 2]])


def hex_boundary_triangles(H):
    """Boundary quads of a hex mesh in the reference's hex_face_table vertex order
    (global_types.h:154-162, kept by build_connectivity for faces of one hex), each split
    into (q0, q1, q2) and (q2, q3, q0) as extract_surface_conforming_mesh does
    (global_functions.cpp:1038-1054). Returns F (2 * n_boundary_quads, 3) in vertex ids of H."""
    Q = H[:, HEX_FACE_TABLE].reshape(-1, 4)
    key = np.sort(Q, axis=1)
    _, inv, cnt = np.unique(key, axis=0, return_inverse=True, return_counts=True)
    B = Q[cnt[inv.ravel()] == 1]
    return np.concatenate([B[:, [0, 1, 2]], B[:, [2, 3, 0]]]).astype(np.int32)


def octree_frame(SV, num_voxels):
    """octree_mesh's frame for a surface (grid_hex_meshing.cpp:463-491, no scaffold layers):
    voxel size (largest bounding-box extent / num_voxels), origin (the box's min corner),
    grid (next power of two of each extent in voxels) and the mesh_transform that centres the
    grid on the box. Returns (voxel_size, origin, grid, mesh_transform)."""
    lo, hi = SV.min(axis=0), SV.max(axis=0)
    vs = float((hi - lo).max()) / num_voxels
    grid = np.array([1 << int(np.ceil(np.log2(np.ceil(e / vs)))) for e in (hi - lo)], dtype=np.int64)
    mt = (lo + hi) / 2 - (lo + (lo + vs * grid.astype(np.float64))) * 0.5
    return vs, lo, grid, mt


def _leaf_keys(x, y, z, level):
    return (x.astype(np.int64) | (y.astype(np.int64) << 16) | (z.astype(np.int64) << 32) |
            (level.astype(np.int64) << 48))


# per face of HEX_FACE_TABLE: the axis it is normal to and whether it lies on the far side
_FACE_AXIS = np.array([2, 2, 1, 0, 1, 0])
_FACE_FAR = np.array([0, 1, 0, 0, 1, 1])


def octree_leaf_boundary(leaves, H):
    """Boundary quads of a hex mesh made of octree leaves (n x 4 rows x, y, z, extent in grid
    units; H their corner node ids in Hybrid::vs order), which may hold hanging nodes where a
    leaf meets leaves of another size. A leaf face is a boundary face when the point just
    outside it at in-face fractions (0.3, 0.4) lies in no leaf of the mesh: with 2:1 graded
    leaves that point is interior to whichever leaf covers it, so faces shared with smaller
    or larger leaves count as inner faces. Returns the boundary quads (m, 4) in
    HEX_FACE_TABLE vertex order."""
    leaves = np.asarray(leaves, dtype=np.int64)
    ext = leaves[:, 3]
    lev = np.round(np.log2(ext)).astype(np.int64)
    keys = np.sort(_leaf_keys(leaves[:, 0], leaves[:, 1], leaves[:, 2], lev))
    levels = np.unique(lev)
    out = []
    for f in range(6):
        a, far = _FACE_AXIS[f], _FACE_FAR[f]
        u, v = [d for d in range(3) if d != a]
        p = leaves[:, :3].astype(np.float64)
        p[:, u] += 0.3 * ext
        p[:, v] += 0.4 * ext
        p[:, a] += ext + 0.25 if far else -0.25
        covered = np.zeros(leaves.shape[0], dtype=bool)
        for L in levels:
            e = float(1 << int(L))
            c = np.floor(p / e) * e
            ok = (c >= 0).all(axis=1)
            k = _leaf_keys(c[:, 0].astype(np.int64), c[:, 1].astype(np.int64), c[:, 2].astype(np.int64),
                           np.full(c.shape[0], L))
            i = np.searchsorted(keys, k)
            hit = ok & (i < keys.size) & (keys[np.minimum(i, keys.size - 1)] == k)
            covered |= hit
        out.append(H[~covered][:, HEX_FACE_TABLE[f]])
    return np.concatenate(out).astype(np.int32)


def box_surface(n):
    """The unit cube [-0.5,0.5]^3 as a closed, outward-oriented triangle mesh: each face an
    n x n quad grid split along one diagonal (12 n^2 triangles, shared edge vertices)."""
    t = np.linspace(-0.5, 0.5, n + 1)
    V, F, index = [], [], {}

    def vid(p):
        key = tuple(np.round(np.asarray(p) * 2 * n).astype(np.int64))
        if key not in index:
            index[key] = len(V)
            V.append(p)
        return index[key]

    for axis in range(3):
        u, v = [d for d in range(3) if d != axis]
        for side in (-0.5, 0.5):
            ids = np.empty((n + 1, n + 1), dtype=np.int64)
            for i in range(n + 1):
                for j in range(n + 1):
                    p = [0.0, 0.0, 0.0]
                    p[axis], p[u], p[v] = side, t[i], t[j]
                    ids[i, j] = vid(p)
            for i in range(n):
                for j in range(n):
                    a, b, c, d = ids[i, j], ids[i + 1, j], ids[i + 1, j + 1], ids[i, j + 1]
                    F.append([a, b, c])
                    F.append([a, c, d])
    V = np.array(V, dtype=np.float64)
    F = np.array(F, dtype=np.int32)
    # outward orientation: the normal points away from the centre
    nrm = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    flip = (nrm * V[F].mean(axis=1)).sum(axis=1) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return V, F


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def box_union_surface(nboxes=8, n=46, seed=0):
    """The union of `nboxes` seeded, randomly rotated boxes (sides U[0.3, 0.8], centres
    U[-0.5, 0.5]^3) as the soup of their closed surfaces (12 n^2 triangles each; ~203k at
    n = 46). The soup is the union for the reference's inside tests: compute_sign counts
    signed +z crossings (voxelization.h:221-272), so a point inside k boxes gets a count of
    -k, negative exactly when it lies in at least one box. Returns V, F and each box's
    (rotation, sides, centre)."""
    rng = np.random.default_rng(seed)
    BV, BF = box_surface(n)
    Vs, Fs, boxes = [], [], []
    off = 0
    for _ in range(nboxes):
        sides = rng.uniform(0.3, 0.8, size=3)
        centre = rng.uniform(-0.5, 0.5, size=3)
        Rm = random_rotation(rng)
        Vs.append((BV * sides) @ Rm.T + centre)
        Fs.append(BF + off)
        off += BV.shape[0]
        boxes.append((Rm, sides, centre))
    return np.concatenate(Vs), np.concatenate(Fs).astype(np.int32), boxes


def surface_samples(V, F, n, seed=0):
    """n seeded points distributed uniformly by area over the triangles (V, F): a triangle
    drawn with probability proportional to its area, then a uniform point in it
    (r1, r2 uniform: sqrt(r1) folds the square onto the triangle)."""
    rng = np.random.default_rng(seed)
    a = 0.5 * np.linalg.norm(np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]]), axis=1)
    t = rng.choice(F.shape[0], size=n, p=a / a.sum())
    r1, r2 = rng.random(n), rng.random(n)
    s = np.sqrt(r1)
    A, B, C = V[F[t, 0]], V[F[t, 1]], V[F[t, 2]]
    return (1 - s)[:, None] * A + (s * (1 - r2))[:, None] * B + (s * r2)[:, None] * C
