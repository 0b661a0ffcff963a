"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU-baseline leg of
bench.py as the checker. The product path never imports this module.
"""
import ctypes as C
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ENERGY = {"ARAP": 0, "LOG_ARAP": 1, "SYMMETRIC_DIRICHLET": 2, "CONFORMAL": 3, "EXP_CONFORMAL": 4,
          "EXP_SYMMETRIC_DIRICHLET": 5}


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        _LIB = C.CDLL(path)
        _LIB.orc_max_step.restype = C.c_double
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class SlimParams(C.Structure):
    _fields_ = [("energy", C.c_int), ("projection", C.c_int), ("glue", C.c_int), ("iters", C.c_int),
                ("weight_opt", C.c_double), ("lam_proj", C.c_double), ("lam_glue", C.c_double),
                ("lam_C", C.c_double), ("lam_T", C.c_double), ("lam_L", C.c_double)]


class SlimConstraints(C.Structure):
    _fields_ = [("nb", C.c_int), ("b", C.POINTER(C.c_int)), ("bc", C.POINTER(C.c_double)),
                ("nranges", C.c_int), ("ranges", C.POINTER(C.c_int)),
                ("nC", C.c_int), ("idsC", C.POINTER(C.c_int)), ("C", C.POINTER(C.c_double)),
                ("nT", C.c_int), ("idsT", C.POINTER(C.c_int)), ("normalT", C.POINTER(C.c_double)),
                ("disT", C.POINTER(C.c_double)),
                ("nL", C.c_int), ("idsL", C.POINTER(C.c_int)), ("axaL", C.POINTER(C.c_double)),
                ("originL", C.POINTER(C.c_double))]


def grad_ref(V, H, hflag=None, cube=True):
    V, H = _d(V), _i(H)
    nH = H.shape[0]
    G = np.zeros((nH * 8, 3, 4))
    hf = None if hflag is None else np.ascontiguousarray(hflag, dtype=np.uint8)
    lib().orc_grad_ref(V.shape[0], _p(V, C.c_double), nH, _p(H, C.c_int),
                       None if hf is None else _p(hf, C.c_uint8), int(cube), _p(G, C.c_double))
    return G


def slim_m_opt(V, H, G, UV, iters, energy=2, weight_opt=1.0, projection=False, lam_proj=1e5,
               lam_C=0.0, lam_T=0.0, lam_L=0.0, b=None, bc=None, idsC=None, Cpos=None, idsT=None,
               normalT=None, disT=None, idsL=None, axaL=None, originL=None):
    """Restated optimization::slim_m_opt. Returns (UV, energy, E_q, E_s, log[iters,7], aL)."""
    V, H, G = _d(V), _i(H), _d(G)
    UV = _d(UV).copy()
    e3 = np.zeros((0, 3))
    e1 = np.zeros(0)
    ei = np.zeros(0, dtype=np.int32)
    b = _i(ei if b is None else b); bc = _d(e3 if bc is None else bc)
    idsC = _i(ei if idsC is None else idsC); Cpos = _d(e3 if Cpos is None else Cpos)
    idsT = _i(ei if idsT is None else idsT); normalT = _d(e3 if normalT is None else normalT)
    disT = _d(e1 if disT is None else disT)
    idsL = _i(ei if idsL is None else idsL); axaL = _d(e3 if axaL is None else axaL)
    originL = _d(e3 if originL is None else originL)
    rng = np.zeros(0, dtype=np.int32)
    prm = SlimParams(energy, int(projection), 0, iters, weight_opt, lam_proj, 0.0, lam_C, lam_T, lam_L)
    cs = SlimConstraints(b.size, _p(b, C.c_int), _p(bc, C.c_double), 0, _p(rng, C.c_int),
                         idsC.size, _p(idsC, C.c_int), _p(Cpos, C.c_double),
                         idsT.size, _p(idsT, C.c_int), _p(normalT, C.c_double), _p(disT, C.c_double),
                         idsL.size, _p(idsL, C.c_int), _p(axaL, C.c_double), _p(originL, C.c_double))
    E, Eq, Es = C.c_double(), C.c_double(), C.c_double()
    log = np.zeros((max(iters, 1), 7))
    aL = np.zeros(max(idsL.size, 1))
    rc = lib().orc_slim_m_opt(V.shape[0], _p(V, C.c_double), H.shape[0], _p(H, C.c_int), _p(G, C.c_double),
                              _p(UV, C.c_double), C.byref(prm), C.byref(cs), C.byref(E), C.byref(Eq),
                              C.byref(Es), _p(log, C.c_double), _p(aL, C.c_double))
    if rc != 0:
        raise RuntimeError("orc_slim_m_opt failed: %d" % rc)
    return UV, E.value, Eq.value, Es.value, log[:iters], aL[:idsL.size]


def local_step(H, G, X, energy=2):
    H, G, X = _i(H), _d(G), _d(X)
    nT = H.shape[0] * 8
    W = np.zeros((nT, 3, 3)); R = np.zeros((nT, 3, 3)); E = np.zeros(nT)
    lib().orc_local_step(X.shape[0], H.shape[0], _p(H, C.c_int), _p(G, C.c_double), _p(X, C.c_double),
                         energy, _p(W, C.c_double), _p(R, C.c_double), _p(E, C.c_double))
    return W, R, E


def svd3(A):
    A = _d(A).reshape(3, 3)
    U = np.zeros((3, 3)); s = np.zeros(3); Vt = np.zeros((3, 3))
    lib().orc_svd3(_p(A, C.c_double), _p(U, C.c_double), _p(s, C.c_double), _p(Vt, C.c_double))
    return U, s, Vt


def max_step(H, X, D):
    H, X, D = _i(H), _d(X), _d(D)
    return lib().orc_max_step(X.shape[0], H.shape[0], _p(H, C.c_int), _p(X, C.c_double), _p(D, C.c_double))


def scaled_jacobian(X, H):
    X, H = _d(X), _i(H)
    mn, avg = C.c_double(), C.c_double()
    per = np.zeros(H.shape[0])
    lib().orc_scaled_jacobian(X.shape[0], _p(X, C.c_double), H.shape[0], _p(H, C.c_int), C.byref(mn),
                              C.byref(avg), _p(per, C.c_double))
    return mn.value, avg.value, per


def signed_distance_pseudonormal(P, V, F, FN, VN, FEN):
    P, V, F, FN, VN, FEN = _d(P), _d(V), _i(F), _d(FN), _d(VN), _d(FEN)
    nq = P.shape[0]
    S = np.zeros(nq); I = np.zeros(nq, dtype=np.int32); Cc = np.zeros((nq, 3)); N = np.zeros((nq, 3))
    sq = np.zeros(nq)
    lib().orc_signed_distance_pseudonormal(nq, _p(P, C.c_double), V.shape[0], _p(V, C.c_double), F.shape[0],
                                           _p(F, C.c_int), _p(FN, C.c_double), _p(VN, C.c_double),
                                           _p(FEN, C.c_double), _p(S, C.c_double), _p(I, C.c_int),
                                           _p(Cc, C.c_double), _p(N, C.c_double), _p(sq, C.c_double))
    return S, I, Cc, N, sq


def cell_facet_overlap(cells, V, F, origin, voxel_size, mesh_transform, stop_extent):
    cells, V, F = _i(cells), _d(V), _i(F)
    o, mt = _d(origin), _d(mesh_transform)
    flag = np.zeros(cells.shape[0], dtype=np.uint8)
    lib().orc_cell_facet_overlap(cells.shape[0], _p(cells, C.c_int), V.shape[0], _p(V, C.c_double), F.shape[0],
                                 _p(F, C.c_int), _p(o, C.c_double), C.c_double(voxel_size), _p(mt, C.c_double),
                                 int(stop_extent), _p(flag, C.c_uint8))
    return flag


def voxel_parity(V, F, nx, ny, nz, origin, spacing):
    V, F, o = _d(V), _i(F), _d(origin)
    out = np.zeros(nx * ny * nz, dtype=np.uint8)
    lib().orc_voxel_parity(V.shape[0], _p(V, C.c_double), F.shape[0], _p(F, C.c_int), nx, ny, nz, _p(o, C.c_double),
                           C.c_double(spacing), _p(out, C.c_uint8))
    return out.reshape(nz, ny, nx)


def cell_parity(V, F, cells, origin, spacing):
    V, F, o, cells = _d(V), _i(F), _d(origin), _i(cells)
    out = np.zeros(cells.shape[0], dtype=np.float32)
    lib().orc_cell_parity(V.shape[0], _p(V, C.c_double), F.shape[0], _p(F, C.c_int), cells.shape[0],
                          _p(cells, C.c_int), _p(o, C.c_double), C.c_double(spacing), _p(out, C.c_float))
    return out


def linear_system(V, H, G, UV, energy=2, weight_opt=1.0, lam_C=0.0, lam_T=0.0, lam_L=0.0, idsC=None, Cpos=None,
                  idsT=None, normalT=None, disT=None, idsL=None, axaL=None, originL=None, dense=True, p=None):
    """Global system of the first SLIM iteration at UV: (L dense or None, rhs), plus (L p, diag)
    when p is given."""
    H, G, UV = _i(H), _d(G), _d(UV)
    e3 = np.zeros((0, 3)); e1 = np.zeros(0); ei = np.zeros(0, dtype=np.int32)
    idsC = _i(ei if idsC is None else idsC); Cpos = _d(e3 if Cpos is None else Cpos)
    idsT = _i(ei if idsT is None else idsT); normalT = _d(e3 if normalT is None else normalT)
    disT = _d(e1 if disT is None else disT)
    idsL = _i(ei if idsL is None else idsL); axaL = _d(e3 if axaL is None else axaL)
    originL = _d(e3 if originL is None else originL)
    prm = SlimParams(energy, 0, 0, 0, weight_opt, 1e5, 0.0, lam_C, lam_T, lam_L)
    cs = SlimConstraints(0, _p(ei, C.c_int), _p(e3, C.c_double), 0, _p(ei, C.c_int),
                         idsC.size, _p(idsC, C.c_int), _p(Cpos, C.c_double),
                         idsT.size, _p(idsT, C.c_int), _p(normalT, C.c_double), _p(disT, C.c_double),
                         idsL.size, _p(idsL, C.c_int), _p(axaL, C.c_double), _p(originL, C.c_double))
    n = 3 * UV.shape[0] + idsL.size
    L = np.zeros((n, n)) if dense else None
    rhs = np.zeros(n)
    pv = None if p is None else _d(p)
    Lp = np.zeros(n); dg = np.zeros(n)
    lib().orc_linear_system(UV.shape[0], H.shape[0], _p(H, C.c_int), _p(G, C.c_double), _p(UV, C.c_double),
                            C.byref(prm), C.byref(cs), None if L is None else _p(L, C.c_double), _p(rhs, C.c_double),
                            None if pv is None else _p(pv, C.c_double), _p(Lp, C.c_double), _p(dg, C.c_double))
    if p is not None:
        return L, rhs, Lp, dg
    return L, rhs


def sd_grad(H, G, X, M=1.0, lam_C=0.0, lam_T=0.0, lam_L=0.0, idsC=None, Cpos=None, idsT=None, normalT=None,
            disT=None, idsL=None, axaL=None, originL=None, aL=None):
    """Energy and gradient of the symmetric-Dirichlet SLIM energy (SVD form)."""
    H, G, X = _i(H), _d(G), _d(X)
    e3 = np.zeros((0, 3)); e1 = np.zeros(0); ei = np.zeros(0, dtype=np.int32)
    idsC = _i(ei if idsC is None else idsC); Cpos = _d(e3 if Cpos is None else Cpos)
    idsT = _i(ei if idsT is None else idsT); normalT = _d(e3 if normalT is None else normalT)
    disT = _d(e1 if disT is None else disT)
    idsL = _i(ei if idsL is None else idsL); axaL = _d(e3 if axaL is None else axaL)
    originL = _d(e3 if originL is None else originL)
    aL = _d(np.zeros(idsL.size) if aL is None else aL)
    cs = SlimConstraints(0, _p(ei, C.c_int), _p(e3, C.c_double), 0, _p(ei, C.c_int),
                         idsC.size, _p(idsC, C.c_int), _p(Cpos, C.c_double),
                         idsT.size, _p(idsT, C.c_int), _p(normalT, C.c_double), _p(disT, C.c_double),
                         idsL.size, _p(idsL, C.c_int), _p(axaL, C.c_double), _p(originL, C.c_double))
    out = np.zeros_like(X)
    E = C.c_double()
    lib().orc_sd_grad(X.shape[0], H.shape[0], _p(H, C.c_int), _p(G, C.c_double), _p(X, C.c_double), C.c_double(M),
                      C.byref(cs), C.c_double(lam_C), C.c_double(lam_T), C.c_double(lam_L), _p(aL, C.c_double),
                      _p(out, C.c_double), C.byref(E))
    return E.value, out


def point_line_projection(A, B, P):
    """Feet of the points P on the segments A-B (all n x 3), as point_line_projection."""
    A, B, P = _d(A), _d(B), _d(P)
    out = np.zeros_like(P)
    lib().orc_point_line_projection(P.shape[0], _p(A, C.c_double), _p(B, C.c_double), _p(P, C.c_double),
                                    _p(out, C.c_double))
    return out


def line_projection(V, curves, circles, P, curve_of):
    """origin_L, Axa_L of LINE vertices P (n x 3) on their curves (lists of vertex ids)."""
    V, P, curve_of = _d(V), _d(P), _i(curve_of)
    off = _i(np.concatenate([[0], np.cumsum([len(c) for c in curves])]))
    cvs = _i(np.concatenate([np.asarray(c, dtype=np.int32) for c in curves]))
    circ = np.ascontiguousarray(circles, dtype=np.uint8)
    n = P.shape[0]
    origin = np.zeros((n, 3)); axis = np.zeros((n, 3))
    lib().orc_line_projection(V.shape[0], _p(V, C.c_double), len(curves), _p(off, C.c_int), _p(cvs, C.c_int),
                              _p(circ, C.c_uint8), n, _p(P, C.c_double), _p(curve_of, C.c_int),
                              _p(origin, C.c_double), _p(axis, C.c_double))
    return origin, axis


def metro_hausdorff(V0, F0, V1, F1):
    """(bbox diagonal, max, mean) of the vertex-sampled metro Hausdorff distance."""
    V0, F0, V1, F1 = _d(V0), _i(F0), _d(V1), _i(F1)
    out = np.zeros(3)
    lib().orc_metro_hausdorff(V0.shape[0], _p(V0, C.c_double), F0.shape[0], _p(F0, C.c_int), V1.shape[0],
                              _p(V1, C.c_double), F1.shape[0], _p(F1, C.c_int), _p(out, C.c_double))
    return tuple(out)
