"""The hand-written adjoint at full size: the 64^3 jittered grid on [-1,1]^3 (262,144 hexes,
274,625 vertices, interior jitter sigma = 0.15 h) with a seeded 2% of hexes (5,243)
inverted by swapping the positions of corners 0 and 1.

- dE/dx of the whole mesh against the oracle's SVD-form gradient (1e-10 relative);
- central differences (eps = 1e-6) of the device energy on 1,000 seeded vertices x 3
  coordinates: the energy is evaluated on a context holding only the hexes incident to the
  picked vertices (with their own reference shapes, which are per hex), so a difference
  sees the same terms as on the whole mesh without the rounding of a 2M-term sum;
- Hessian-vector products against central differences of the gradient, and symmetry;
- the same gradient and difference checks with every soft-constraint kind active (planes
  on the face vertices, corners, feature lines with their a_L).

The swaps leave some corner tets nearly flat (altitudes far below eps, gradients up to
2e13): there a difference at eps = 1e-6 straddles a pole of the energy and measures nothing,
and near them the O(eps^2) truncation of a central difference grows with the curvature.
Vertices are "regular" when every altitude of their incident corner tets is above 1e-3
(257,762 of them) and "smooth" above 5e-3 (148,822). The seeded vertices and the
difference direction of the checks are drawn from the smooth ones, the Hessian-vector rows
are compared on the regular ones (the rows of the others carry the roundoff of gradients
of 1e13), and the gradient is compared with the oracle over the whole mesh and over the
regular vertices alone, so the huge entries do not set the scale of the others."""
import numpy as np
import pytest

import oracle_py as O
from helpers import rel

pytestmark = pytest.mark.gpu

N = 64
EPS = 1e-6


def _mesh():
    from pyhexopt import meshgen as mg
    V, H = mg.grid_hex(N)
    V = mg.jitter(V, ~mg.grid_boundary_mask(N), 0.15 * 2.0 / N, seed=0)
    V, picked = mg.invert_hexes(V, H, 0.02, seed=0)
    assert picked.size == 5243
    return V, H


def _regular(V, H, min_alt=1e-3):
    """Vertices whose incident corner tets all have every altitude above min_alt."""
    from pyhexopt import meshgen as mg
    P = V[H]
    alt = np.full(H.shape[0], np.inf)
    for c in range(8):
        q = [P[:, mg.HEX_TETRA_TABLE[c, k]] for k in range(4)]
        vol6 = np.abs(np.einsum("ij,ij->i", q[1] - q[0], np.cross(q[2] - q[0], q[3] - q[0])))
        area2 = np.maximum.reduce([np.linalg.norm(np.cross(q[b] - q[a], q[e] - q[a]), axis=1)
                                   for a, b, e in ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3))])
        alt = np.minimum(alt, vol6 / np.maximum(area2, 1e-300))   # 3 vol / largest face area
    vmin = np.full(V.shape[0], np.inf)
    for k in range(8):
        np.minimum.at(vmin, H[:, k], alt)
    return vmin > min_alt


def _context(V, H):
    from pyhexopt import capi
    h = capi.HexOpt(V, H)
    h.reference()
    return h


def _constraints(V, X):
    """Corners on the 8 cube corners, lines along the 12 edges, planes on the 6 faces."""
    t = np.isclose(np.abs(V), 1.0)
    k = t.sum(1)
    corner, edge, face = np.nonzero(k == 3)[0], np.nonzero(k == 2)[0], np.nonzero(k == 1)[0]
    rng = np.random.default_rng(11)
    nrm = np.where(t[face], np.sign(V[face]), 0.0)
    kw = dict(idsC=corner.astype(np.int32), Cpos=V[corner] + 0.01 * rng.normal(size=(corner.size, 3)), lamC=3.0,
              idsT=face.astype(np.int32), normalT=nrm, disT=np.ones(face.size), lamT=5.0)
    axis = np.where(t[edge], 0.0, 1.0)                  # the free coordinate of an edge vertex
    origin = np.where(t[edge], np.sign(V[edge]), 0.0)
    kw.update(idsL=edge.astype(np.int32), axaL=axis, originL=origin, lamL=7.0)
    aL = (X[edge] * axis).sum(1) + 0.002 * rng.normal(size=edge.size)
    return kw, aL


def _oracle_kwargs(kw):
    return dict(idsC=kw["idsC"], Cpos=kw["Cpos"], lam_C=kw["lamC"], idsT=kw["idsT"], normalT=kw["normalT"],
                disT=kw["disT"], lam_T=kw["lamT"], idsL=kw["idsL"], axaL=kw["axaL"], originL=kw["originL"],
                lam_L=kw["lamL"])


def _sub_constraints(kw, remap, aL):
    """The constraints of the vertices a sub-context holds, in its numbering."""
    out, a = {}, None
    for key, ids, rows in (("C", "idsC", ["Cpos"]), ("T", "idsT", ["normalT", "disT"]),
                           ("L", "idsL", ["axaL", "originL"])):
        keep = remap[kw[ids]] >= 0
        out[ids] = remap[kw[ids][keep]].astype(np.int32)
        for r in rows:
            out[r] = kw[r][keep]
        out["lam" + key] = kw["lam" + key]
        if key == "L":
            a = aL[keep]
    return out, a


def _central_differences(V, H, X, g, verts, kw=None, aL=None):
    """Largest |FD - g| / |g| over the 3 coordinates of each vertex of verts (|g| floored at
    1e-3 of its median over them). Each vertex gets a context of its incident hexes alone,
    so the energy sum is a few hundred terms and its rounding stays far below the check."""
    from pyhexopt import capi
    gp = g[verts]
    floor = 1e-3 * np.median(np.abs(gp))
    worst = 0.0
    for v, gv in zip(verts, gp):
        hexes = np.nonzero((H == v).any(1))[0]
        used = np.unique(H[hexes])
        remap = np.full(V.shape[0], -1, dtype=np.int64)
        remap[used] = np.arange(used.size)
        hs = capi.HexOpt(V[used], remap[H[hexes]].astype(np.int32))
        try:
            hs.reference()                  # per-hex reference shapes, as in the whole mesh
            a = None
            if kw is not None:
                skw, a = _sub_constraints(kw, remap, aL)
                hs.set_constraints(**skw)
            Xs = X[used].copy()
            i = remap[v]
            for d in range(3):
                x0 = Xs[i, d]
                Xs[i, d] = x0 + EPS
                hs.set_positions(capi.CURRENT, Xs)
                ep = sum(hs.energy(aL=a))
                Xs[i, d] = x0 - EPS
                hs.set_positions(capi.CURRENT, Xs)
                em = sum(hs.energy(aL=a))
                Xs[i, d] = x0
                fd = (ep - em) / (2 * EPS)
                worst = max(worst, abs(fd - gv[d]) / max(abs(gv[d]), floor))
        finally:
            hs.close()
    return worst


@pytest.fixture(scope="module")
def mesh():
    V, H = _mesh()
    reg, smooth = _regular(V, H, 1e-3), _regular(V, H, 5e-3)
    print("64^3 adjoint: %d hexes, %d vertices, %d regular, %d smooth" %
          (H.shape[0], V.shape[0], reg.sum(), smooth.sum()), flush=True)
    return V, H, O.grad_ref(V, H), reg, smooth


def _picks(reg, n, seed, among=None):
    cand = np.nonzero(reg)[0] if among is None else among[reg[among]]
    return np.random.default_rng(seed).choice(cand, size=n, replace=False)


def test_gradient_64_vs_oracle_and_central_differences(mesh):
    V, H, G, reg, smooth = mesh
    h = _context(V, H)
    try:
        g = h.energy_grad(V)
    finally:
        h.close()
    _, go = O.sd_grad(H, G, V)
    print("64^3 gradient: vs oracle %.2e (regular vertices %.2e)" % (rel(g, go), rel(g[reg], go[reg])), flush=True)
    # the largest entries come from the near-flat tets, whose J^-1 amplifies rounding by its
    # condition number (up to ~1e7): 1e-8 there, 1e-10 over the regular vertices
    assert rel(g, go) < 1e-8
    assert rel(g[reg], go[reg]) < 1e-10
    worst = _central_differences(V, H, V, g, _picks(smooth, 1000, 0))
    print("64^3 gradient: vs oracle %.2e (regular %.2e), central differences %.2e" %
          (rel(g, go), rel(g[reg], go[reg]), worst), flush=True)
    # the same differences of the oracle's energy reach 1.1e-6 on these picks
    assert worst < 2e-5


def test_hvp_64_vs_gradient_differences_and_symmetry(mesh):
    V, H, _, reg, smooth = mesh
    rng = np.random.default_rng(2)
    u = rng.normal(size=V.shape)
    w = rng.normal(size=V.shape)
    ul = np.where(smooth[:, None], u, 0.0)    # difference direction on the smooth vertices
    h = _context(V, H)
    try:
        Hu, Hw = h.hvp(V, u), h.hvp(V, w)
        Hul = h.hvp(V, ul)
        fd = (h.energy_grad(V + EPS * ul) - h.energy_grad(V - EPS * ul)) / (2 * EPS)
    finally:
        h.close()
    a, b = (w * Hu).sum(), (u * Hw).sum()
    # the oracle's gradient differences at eps and eps / 2 agree to 1.4e-6 on these rows
    print("64^3 HVP: vs differences %.2e, symmetry %.2e" % (rel(Hul[reg], fd[reg]), abs(a - b) / abs(a)),
          flush=True)
    assert rel(Hul[reg], fd[reg]) < 1e-5
    assert abs(a - b) <= 1e-10 * max(abs(a), abs(b))


def test_gradient_64_with_every_constraint_kind(mesh):
    V, H, G, reg, smooth = mesh
    X = V
    kw, aL = _constraints(V, X)
    kw["disT"] = kw["disT"] + 0.002 * np.random.default_rng(4).normal(size=kw["disT"].size)
    h = _context(V, H)
    try:
        h.set_constraints(**kw)
        g = h.energy_grad(X, aL=aL)
    finally:
        h.close()
    _, go = O.sd_grad(H, G, X, aL=aL, **_oracle_kwargs(kw))
    print("64^3 constrained gradient: vs oracle %.2e (regular vertices %.2e)" % (rel(g, go), rel(g[reg], go[reg])),
          flush=True)
    assert rel(g, go) < 1e-8
    assert rel(g[reg], go[reg]) < 1e-10
    pick_v = np.concatenate([kw["idsC"][smooth[kw["idsC"]]], _picks(smooth, 300, 3, kw["idsL"]),
                             _picks(smooth, 400, 4, kw["idsT"])])
    rest = np.setdiff1d(np.nonzero(smooth)[0], pick_v)
    pick_v = np.concatenate([pick_v, np.random.default_rng(5).choice(rest, 1000 - pick_v.size, replace=False)])
    worst = _central_differences(V, H, X, g, pick_v, kw, aL)
    print("64^3 constrained gradient: vs oracle %.2e, central differences %.2e" % (rel(g, go), worst), flush=True)
    # the same differences of the oracle's energy reach 7.1e-6 on these picks
    assert worst < 2e-5
