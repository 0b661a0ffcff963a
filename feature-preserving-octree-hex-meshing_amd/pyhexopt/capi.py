"""ctypes binding of libhexopt.so (include/hexopt.h).

The library must have been built (feature-preserving-octree-hex-meshing_amd/Makefile); there is
no fallback: a missing library or a missing GPU raises.
"""
import ctypes as C
import os
import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "libhexopt.so")

ARAP, LOG_ARAP, SYMMETRIC_DIRICHLET, CONFORMAL, EXP_CONFORMAL, EXP_SYMMETRIC_DIRICHLET = range(6)
REST, CURRENT = 0, 1

EXPORTS = ["hexopt_create", "hexopt_create_error", "hexopt_destroy", "hexopt_last_error", "hexopt_set_positions",
           "hexopt_get_positions", "hexopt_device_positions", "hexopt_reference", "hexopt_set_constraints",
           "hexopt_slim_solve", "hexopt_energy", "hexopt_energy_grad", "hexopt_hvp", "hexopt_scaled_jacobian",
           "hexopt_synchronize", "hexopt_layout_info", "hexopt_debug_system", "hexopt_opt32_solve",
           "hexopt_time_kernel", "hexopt_copy_positions", "hexopt_surface_create", "hexopt_surface_error",
           "hexopt_surface_last_error", "hexopt_surface_destroy", "hexopt_surface_normals",
           "hexopt_compute_surface_normals", "hexopt_surface_closest", "hexopt_project_planes",
           "hexopt_occ_classify_cells", "hexopt_occ_voxel_parity", "hexopt_occ_cell_parity", "hexopt_hex_inside",
           "hexopt_curves_create", "hexopt_curves_error", "hexopt_curves_destroy", "hexopt_project_lines",
           "hexopt_refresh_lines", "hexopt_metro_hausdorff", "hexopt_glue_setup", "hexopt_glue_solve",
           "hexopt_octree_build", "hexopt_octree_error", "hexopt_octree_destroy", "hexopt_octree_sizes",
           "hexopt_octree_copy", "hexopt_octree_refine", "hexopt_scaled_jacobian_noalloc", "hexopt_noalloc_error", "hexopt_project_patches", "hexopt_debug_operator"]

_LIB = None


class Params(C.Structure):
    _fields_ = [("energy", C.c_int), ("weight_opt", C.c_double), ("exp_factor", C.c_double),
                ("projection", C.c_int), ("lamda_projection", C.c_double), ("tet_mass", C.c_double)]


class Constraints(C.Structure):
    _fields_ = [("nb", C.c_int), ("b", C.POINTER(C.c_int32)), ("bc", C.POINTER(C.c_double)),
                ("nC", C.c_int), ("ids_C", C.POINTER(C.c_int32)), ("C", C.POINTER(C.c_double)), ("lamda_C", C.c_double),
                ("nT", C.c_int), ("ids_T", C.POINTER(C.c_int32)), ("normal_T", C.POINTER(C.c_double)),
                ("dis_T", C.POINTER(C.c_double)), ("lamda_T", C.c_double),
                ("nL", C.c_int), ("ids_L", C.POINTER(C.c_int32)), ("Axa_L", C.POINTER(C.c_double)),
                ("origin_L", C.POINTER(C.c_double)), ("lamda_L", C.c_double)]


class IterLog(C.Structure):
    _fields_ = [("energy", C.c_double), ("e_quality", C.c_double), ("e_soft", C.c_double),
                ("step_bound", C.c_double), ("step", C.c_double), ("cg_iters", C.c_int), ("cg_error", C.c_double)]


class GlueLog(C.Structure):
    _fields_ = [("energy", C.c_double), ("e_quality", C.c_double), ("e_soft", C.c_double),
                ("lamda_glue", C.c_double), ("min_sj", C.c_double), ("avg_sj", C.c_double), ("step", C.c_double),
                ("cg_iters", C.c_int)]


class Opt32Options(C.Structure):
    _fields_ = [("cg_iters", C.c_int), ("refine_iters", C.c_int), ("residual_every", C.c_int)]


class Opt32Log(C.Structure):
    _fields_ = [("energy", C.c_double), ("e_quality", C.c_double), ("e_soft", C.c_double),
                ("step_bound", C.c_double), ("step", C.c_double), ("cg_iters", C.c_int), ("cg_error", C.c_double),
                ("residual64", C.c_double), ("inverted_tets", C.c_int), ("residual64_refined", C.c_double)]


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libhexopt.so is not built: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.hexopt_create.restype = C.c_void_p
        L.hexopt_create.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int]
        L.hexopt_create_error.restype = C.c_char_p
        L.hexopt_last_error.restype = C.c_char_p
        L.hexopt_last_error.argtypes = [C.c_void_p]
        L.hexopt_destroy.argtypes = [C.c_void_p]
        L.hexopt_device_positions.restype = C.c_void_p
        L.hexopt_device_positions.argtypes = [C.c_void_p, C.c_int]
        for name in ["hexopt_set_positions", "hexopt_get_positions", "hexopt_reference", "hexopt_set_constraints",
                     "hexopt_slim_solve", "hexopt_energy", "hexopt_energy_grad", "hexopt_hvp",
                     "hexopt_scaled_jacobian", "hexopt_synchronize", "hexopt_layout_info", "hexopt_debug_system",
                     "hexopt_opt32_solve", "hexopt_time_kernel", "hexopt_copy_positions",
                     "hexopt_surface_normals", "hexopt_compute_surface_normals", "hexopt_surface_closest",
                     "hexopt_project_planes", "hexopt_occ_classify_cells", "hexopt_occ_voxel_parity",
                     "hexopt_occ_cell_parity", "hexopt_hex_inside", "hexopt_project_lines",
                     "hexopt_refresh_lines", "hexopt_metro_hausdorff", "hexopt_glue_setup", "hexopt_glue_solve"]:
            getattr(L, name).restype = C.c_int
        L.hexopt_octree_build.restype = C.c_void_p
        L.hexopt_octree_build.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_int,
                                          C.c_int, C.c_int]
        L.hexopt_octree_error.restype = C.c_char_p
        L.hexopt_octree_destroy.argtypes = [C.c_void_p]
        L.hexopt_octree_destroy.restype = None
        L.hexopt_octree_sizes.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.hexopt_octree_sizes.restype = C.c_int
        L.hexopt_octree_copy.argtypes = [C.c_void_p] * 5
        L.hexopt_octree_copy.restype = C.c_int
        L.hexopt_octree_refine.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.hexopt_octree_refine.restype = C.c_int
        L.hexopt_curves_create.restype = C.c_void_p
        L.hexopt_curves_create.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.hexopt_curves_error.restype = C.c_char_p
        L.hexopt_curves_destroy.argtypes = [C.c_void_p]
        L.hexopt_surface_create.restype = C.c_void_p
        L.hexopt_surface_create.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_int]
        L.hexopt_surface_error.restype = C.c_char_p
        L.hexopt_surface_last_error.restype = C.c_char_p
        L.hexopt_surface_last_error.argtypes = [C.c_void_p]
        L.hexopt_surface_destroy.argtypes = [C.c_void_p]
        L.hexopt_scaled_jacobian_noalloc.restype = C.c_int
        L.hexopt_scaled_jacobian_noalloc.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                                     C.c_void_p, C.c_void_p]
        L.hexopt_noalloc_error.restype = C.c_char_p
        L.hexopt_debug_operator.restype = C.c_int
        L.hexopt_project_patches.restype = C.c_int
        L.hexopt_project_patches.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_double]
        _LIB = L
    return _LIB


def _ptr(a, t=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class HexOpt:
    """One context of libhexopt on one GPU (see include/hexopt.h)."""

    def __init__(self, V, H, device=0):
        self.V0 = _d(V)
        self.H = _i(H)
        self.nV, self.nH = self.V0.shape[0], self.H.shape[0]
        self._keep = []
        self.h = lib().hexopt_create(self.nV, self.V0.ctypes.data, self.nH, self.H.ctypes.data, device)
        if not self.h:
            raise RuntimeError("hexopt_create: " + lib().hexopt_create_error().decode())

    def close(self):
        if self.h:
            lib().hexopt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            raise RuntimeError("%s failed (%d): %s" % (what, rc, lib().hexopt_last_error(self.h).decode()))

    def set_positions(self, which, V):
        V = _d(V)
        assert V.shape == (self.nV, 3)
        self._chk(lib().hexopt_set_positions(C.c_void_p(self.h), which, _ptr(V)), "set_positions")

    def get_positions(self, which=CURRENT):
        out = np.zeros((self.nV, 3))
        self._chk(lib().hexopt_get_positions(C.c_void_p(self.h), which, _ptr(out)), "get_positions")
        return out

    def rest_from_current(self):
        self._chk(lib().hexopt_copy_positions(C.c_void_p(self.h), CURRENT, REST), "copy_positions")

    def reference(self, hflag=None, cube=True):
        hf = None if hflag is None else np.ascontiguousarray(hflag, dtype=np.uint8)
        self._chk(lib().hexopt_reference(C.c_void_p(self.h), _ptr(hf, C.c_uint8), int(cube)), "reference")

    def set_constraints(self, b=None, bc=None, idsC=None, Cpos=None, lamC=0.0, idsT=None, normalT=None, disT=None,
                        lamT=0.0, idsL=None, axaL=None, originL=None, lamL=0.0):
        def ids(a):
            return _i(np.zeros(0) if a is None else a)

        def arr(a, n):
            return _d(np.zeros((n, 3)) if a is None else a)

        b, idsC, idsT, idsL = ids(b), ids(idsC), ids(idsT), ids(idsL)
        bc, Cpos = arr(bc, b.size), arr(Cpos, idsC.size)
        normalT, axaL, originL = arr(normalT, idsT.size), arr(axaL, idsL.size), arr(originL, idsL.size)
        disT = _d(np.zeros(idsT.size) if disT is None else disT)
        keep = [b, bc, idsC, Cpos, idsT, normalT, disT, idsL, axaL, originL]
        self._keep = keep
        cs = Constraints(b.size, _ptr(b, C.c_int32), _ptr(bc), idsC.size, _ptr(idsC, C.c_int32), _ptr(Cpos), lamC,
                         idsT.size, _ptr(idsT, C.c_int32), _ptr(normalT), _ptr(disT), lamT,
                         idsL.size, _ptr(idsL, C.c_int32), _ptr(axaL), _ptr(originL), lamL)
        self.nL = idsL.size
        self._chk(lib().hexopt_set_constraints(C.c_void_p(self.h), C.byref(cs)), "set_constraints")

    @staticmethod
    def params(energy=SYMMETRIC_DIRICHLET, weight_opt=1.0, exp_factor=1.0, projection=False, lam_proj=1e5,
               tet_mass=0.0):
        return Params(energy, weight_opt, exp_factor, int(projection), lam_proj, tet_mass)

    def slim_solve(self, iters, energy=SYMMETRIC_DIRICHLET, weight_opt=1.0, tet_mass=0.0):
        prm = self.params(energy, weight_opt, tet_mass=tet_mass)
        logs = (IterLog * max(iters, 1))()
        E, Eq, Es = C.c_double(), C.c_double(), C.c_double()
        aL = np.zeros(max(getattr(self, "nL", 0), 1))
        self._chk(lib().hexopt_slim_solve(C.c_void_p(self.h), C.byref(prm), iters, logs, C.byref(E), C.byref(Eq),
                                          C.byref(Es), _ptr(aL)), "slim_solve")
        log = np.array([[l.energy, l.e_quality, l.e_soft, l.step_bound, l.step, l.cg_iters, l.cg_error]
                        for l in logs[:iters]]).reshape(iters, 7)
        return E.value, Eq.value, Es.value, log, aL[:getattr(self, "nL", 0)]

    def glue_setup(self, V_orig, H_orig, orig_of, fixed_orig):
        """Exploded-mesh glue (include/hexopt.h): this context holds the exploded mesh."""
        V_orig, H_orig = _d(V_orig), _i(H_orig)
        orig_of = _i(orig_of)
        fx = np.ascontiguousarray(fixed_orig, dtype=np.uint8)
        self.nO = V_orig.shape[0]
        self._chk(lib().hexopt_glue_setup(C.c_void_p(self.h), self.nO, _ptr(V_orig), _ptr(H_orig, C.c_int32),
                                          _ptr(orig_of, C.c_int32), _ptr(fx, C.c_uint8)), "glue_setup")

    def glue_solve(self, max_iters=50, energy=CONFORMAL, weight_opt=1.0):
        """Returns (glued V, log rows [energy, e_quality, e_soft, lamda_glue, min_sj, avg_sj, step,
        cg_iters], iterations, min scaled Jacobian)."""
        prm = self.params(energy, weight_opt)
        logs = (GlueLog * max(max_iters, 1))()
        it, mj = C.c_int(), C.c_double()
        Vg = np.zeros((self.nO, 3))
        self._chk(lib().hexopt_glue_solve(C.c_void_p(self.h), C.byref(prm), max_iters, logs, C.byref(it), _ptr(Vg),
                                          C.byref(mj)), "glue_solve")
        n = it.value
        log = np.array([[l.energy, l.e_quality, l.e_soft, l.lamda_glue, l.min_sj, l.avg_sj, l.step, l.cg_iters]
                        for l in logs[:n]]).reshape(n, 8)
        return Vg, log, n, mj.value

    def energy(self, energy=SYMMETRIC_DIRICHLET, weight_opt=1.0, aL=None):
        prm = self.params(energy, weight_opt)
        Eq, Es = C.c_double(), C.c_double()
        a = None if aL is None else _d(aL)
        self._chk(lib().hexopt_energy(C.c_void_p(self.h), C.byref(prm), _ptr(a), C.byref(Eq), C.byref(Es)), "energy")
        return Eq.value, Es.value

    def energy_grad(self, X, precision=64, weight_opt=1.0, aL=None):
        dt = np.float64 if precision == 64 else np.float32
        X = np.ascontiguousarray(X, dtype=dt)
        G = np.zeros((self.nV, 3), dtype=dt)
        prm = self.params(SYMMETRIC_DIRICHLET, weight_opt)
        a = None if aL is None else _d(aL)
        self._chk(lib().hexopt_energy_grad(C.c_void_p(self.h), C.byref(prm), precision, X.ctypes.data_as(C.c_void_p),
                                           _ptr(a), G.ctypes.data_as(C.c_void_p)), "energy_grad")
        return G

    def hvp(self, X, v, precision=64, weight_opt=1.0):
        dt = np.float64 if precision == 64 else np.float32
        X = np.ascontiguousarray(X, dtype=dt)
        v = np.ascontiguousarray(v, dtype=dt)
        out = np.zeros((self.nV, 3), dtype=dt)
        prm = self.params(SYMMETRIC_DIRICHLET, weight_opt)
        self._chk(lib().hexopt_hvp(C.c_void_p(self.h), C.byref(prm), precision, X.ctypes.data_as(C.c_void_p),
                                   v.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p)), "hvp")
        return out

    def scaled_jacobian(self, which=CURRENT, per_hex=False):
        mn, avg = C.c_double(), C.c_double()
        ph = np.zeros(self.nH) if per_hex else None
        self._chk(lib().hexopt_scaled_jacobian(C.c_void_p(self.h), which, C.byref(mn), C.byref(avg), _ptr(ph)),
                  "scaled_jacobian")
        return (mn.value, avg.value, ph) if per_hex else (mn.value, avg.value)

    def debug_system(self, p, energy=SYMMETRIC_DIRICHLET, weight_opt=1.0):
        """(rhs, inverse Jacobi diagonal, L p) of the next SLIM iteration at CURRENT, flattened
        [x y z a_L] in caller numbering."""
        prm = self.params(energy, weight_opt)
        p = _d(p)
        rhs, invd, Lp = np.zeros_like(p), np.zeros_like(p), np.zeros_like(p)
        self._chk(lib().hexopt_debug_system(C.c_void_p(self.h), C.byref(prm), _ptr(p), _ptr(rhs), _ptr(invd),
                                            _ptr(Lp)), "debug_system")
        return rhs, invd, Lp

    def debug_operator(self, precision=64, weight_opt=1.0):
        """Per-hex, per-corner-tet SLIM weight blocks (nH, 8, 6) from the FP64 or FP32 local step."""
        prm = self.params(SYMMETRIC_DIRICHLET, weight_opt)
        K = np.zeros((self.nH, 8, 6))
        self._chk(lib().hexopt_debug_operator(C.c_void_p(self.h), C.byref(prm), precision, _ptr(K)),
                  "debug_operator")
        return K

    def opt32_solve(self, iters, cg_iters=20, refine_iters=0, weight_opt=1.0, residual_every=1):
        """FP32 + adjoint optimiser iterations from CURRENT (include/hexopt.h)."""
        prm = self.params(SYMMETRIC_DIRICHLET, weight_opt)
        opt = Opt32Options(cg_iters, refine_iters, residual_every)
        logs = (Opt32Log * max(iters, 1))()
        E = C.c_double()
        self._chk(lib().hexopt_opt32_solve(C.c_void_p(self.h), C.byref(prm), C.byref(opt), iters, logs, C.byref(E)),
                  "opt32_solve")
        cols = ["energy", "e_quality", "e_soft", "step_bound", "step", "cg_iters", "cg_error", "residual64",
                "inverted_tets", "residual64_refined"]
        log = np.array([[getattr(l, k) for k in cols] for l in logs[:iters]]).reshape(iters, len(cols))
        return E.value, log

    def project_planes(self, surface, lamT):
        self._chk(lib().hexopt_project_planes(C.c_void_p(self.h), C.c_void_p(surface.h), C.c_double(lamT)),
                  "project_planes")

    def project_patches(self, patches, patch_of_T, lamT):
        """Tangent planes of the plane vertices from their patches' surfaces (a list of
        Surface); patch_of_T indexes that list, one entry per ids_T row, or is None to reuse
        the previous call's assignment."""
        arr = (C.c_void_p * len(patches))(*[p.h for p in patches])
        pt = None if patch_of_T is None else _i(patch_of_T)
        self._chk(lib().hexopt_project_patches(C.c_void_p(self.h), len(patches), arr, _ptr(pt, C.c_int32),
                                               C.c_double(lamT)), "project_patches")

    def project_lines(self, curves, ids, curve_of, which=CURRENT):
        """origin_L and Axa_L (n x 3) of LINE vertices ids on curves curve_of."""
        ids, curve_of = _i(ids), _i(curve_of)
        n = ids.size
        origin = np.zeros((n, 3)); axis = np.zeros((n, 3))
        self._chk(lib().hexopt_project_lines(C.c_void_p(self.h), C.c_void_p(curves.h), which, n, _ptr(ids, C.c_int32),
                                             _ptr(curve_of, C.c_int32), _ptr(origin), _ptr(axis)), "project_lines")
        return origin, axis

    def refresh_lines(self, curves, curve_of, lamL):
        """Re-project the line constraints onto their curves with weight lamL, on the device."""
        curve_of = _i(curve_of)
        self._chk(lib().hexopt_refresh_lines(C.c_void_p(self.h), C.c_void_p(curves.h), _ptr(curve_of, C.c_int32),
                                             C.c_double(lamL)), "refresh_lines")

    def hex_inside(self, surface, which=CURRENT):
        """Signed distance of every hex's bounding-box centre and the inside flags."""
        S = np.zeros(self.nH)
        flag = np.zeros(self.nH, dtype=np.uint8)
        self._chk(lib().hexopt_hex_inside(C.c_void_p(self.h), C.c_void_p(surface.h), which, _ptr(S),
                                          _ptr(flag, C.c_uint8)), "hex_inside")
        return S, flag

    def time_kernel(self, which, reps=20):
        ms = C.c_double()
        self._chk(lib().hexopt_time_kernel(C.c_void_p(self.h), which, reps, C.byref(ms)), "time_kernel")
        return ms.value

    def synchronize(self):
        self._chk(lib().hexopt_synchronize(C.c_void_p(self.h)), "synchronize")

    def layout_info(self):
        out = np.zeros(6, dtype=np.int32)
        self._chk(lib().hexopt_layout_info(C.c_void_p(self.h), _ptr(out, C.c_int32)), "layout_info")
        return dict(zip(["tiles", "shared", "slots", "max_colours", "ref_mode", "corner_tiles"], out.tolist()))


def scaled_jacobian_noalloc(device, nV, X_ptr, nH, conn_ptr, per_hex_ptr, min_sum_ptr, stream=None):
    """hexopt_scaled_jacobian_noalloc on caller-owned device arrays (integer device addresses:
    X as x[d*nV+v], conn as conn[k*nH+h], per_hex nullable, min_sum two doubles). Queues the
    work on `stream` (an integer hipStream_t handle, None for the default stream)."""
    rc = lib().hexopt_scaled_jacobian_noalloc(device, nV, C.c_void_p(X_ptr), nH, C.c_void_p(conn_ptr),
                                              C.c_void_p(per_hex_ptr) if per_hex_ptr else None,
                                              C.c_void_p(min_sum_ptr), C.c_void_p(stream) if stream else None)
    if rc != 0:
        raise RuntimeError("hexopt_scaled_jacobian_noalloc failed (%d): %s" %
                           (rc, lib().hexopt_noalloc_error().decode()))


def surface_normals(V, F):
    """Face, angle-weighted vertex and per-face-edge normals (host only)."""
    V, F = _d(V), _i(F)
    FN = np.zeros((F.shape[0], 3)); VN = np.zeros((V.shape[0], 3)); FEN = np.zeros((F.shape[0], 9))
    rc = lib().hexopt_compute_surface_normals(V.shape[0], _ptr(V), F.shape[0], _ptr(F, C.c_int32), _ptr(FN), _ptr(VN),
                                              _ptr(FEN))
    if rc != 0:
        raise RuntimeError("hexopt_compute_surface_normals failed (%d)" % rc)
    return FN, VN, FEN


def metro_hausdorff(V0, F0, V1, F1, device=0):
    """(bbox diagonal, max, mean) of the vertex-sampled metro Hausdorff distance."""
    V0, F0, V1, F1 = _d(V0), _i(F0), _d(V1), _i(F1)
    D, mx, mean = C.c_double(), C.c_double(), C.c_double()
    rc = lib().hexopt_metro_hausdorff(V0.shape[0], _ptr(V0), F0.shape[0], _ptr(F0, C.c_int32), V1.shape[0], _ptr(V1),
                                      F1.shape[0], _ptr(F1, C.c_int32), device, C.byref(D), C.byref(mx),
                                      C.byref(mean))
    if rc != 0:
        raise RuntimeError("hexopt_metro_hausdorff failed (%d): %s" % (rc, lib().hexopt_surface_error().decode()))
    return D.value, mx.value, mean.value


class Surface:
    """Input triangle surface for closest-point queries (include/hexopt.h)."""

    def __init__(self, V, F, FN=None, VN=None, FEN=None, device=0):
        self.V, self.F = _d(V), _i(F)
        arrs = [None if a is None else _d(a) for a in (FN, VN, FEN)]
        self._keep = arrs
        ptrs = [None if a is None else a.ctypes.data for a in arrs]
        self.h = lib().hexopt_surface_create(self.V.shape[0], self.V.ctypes.data, self.F.shape[0], self.F.ctypes.data,
                                             ptrs[0], ptrs[1], ptrs[2], device)
        if not self.h:
            raise RuntimeError("hexopt_surface_create: " + lib().hexopt_surface_error().decode())

    def close(self):
        if self.h:
            lib().hexopt_surface_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def closest(self, P):
        P = _d(P)
        nq = P.shape[0]
        S = np.zeros(nq); I = np.zeros(nq, dtype=np.int32); Cc = np.zeros((nq, 3)); N = np.zeros((nq, 3))
        rc = lib().hexopt_surface_closest(C.c_void_p(self.h), nq, _ptr(P), _ptr(S), _ptr(I, C.c_int32), _ptr(Cc),
                                          _ptr(N))
        if rc != 0:
            raise RuntimeError("hexopt_surface_closest failed (%d): %s" %
                               (rc, lib().hexopt_surface_last_error(C.c_void_p(self.h)).decode()))
        return S, I, Cc, N

    def _schk(self, rc, what):
        if rc != 0:
            raise RuntimeError("%s failed (%d): %s" % (what, rc, lib().hexopt_surface_last_error(C.c_void_p(self.h)).decode()))

    def classify_cells(self, cells, origin, voxel_size, mesh_transform=None, stop_extent=1):
        """Subdivision predicate for n x 4 cells (x, y, z, extent): uint8 flags."""
        cells = _i(cells).reshape(-1, 4)
        out = np.zeros(cells.shape[0], dtype=np.uint8)
        mt = None if mesh_transform is None else _d(mesh_transform)
        self._schk(lib().hexopt_occ_classify_cells(C.c_void_p(self.h), cells.shape[0], _ptr(cells, C.c_int32),
                                                    _ptr(_d(origin)), C.c_double(voxel_size), _ptr(mt),
                                                    int(stop_extent), _ptr(out, C.c_uint8)), "occ_classify_cells")
        return out

    def octree(self, grid, origin, voxel_size, mesh_transform=None, stop_extent=1, graded=True, paired=True):
        """OctreeGrid(grid) + subdivide(should_subdivide, graded, paired) + octree_mesh's export:
        dict(leaves (n, 4), nodes (m, 3) integer positions, V (m, 3), H (n, 8), levels,
        evaluated, ms)."""
        o = Octree(self, grid, origin, voxel_size, mesh_transform, stop_extent, graded, paired)
        try:
            return o.result()
        finally:
            o.close()

    def voxel_parity(self, nx, ny, nz, origin, spacing):
        """Inside flags of voxel centres, shaped (nz, ny, nx)."""
        out = np.zeros(nx * ny * nz, dtype=np.uint8)
        self._schk(lib().hexopt_occ_voxel_parity(C.c_void_p(self.h), nx, ny, nz, _ptr(_d(origin)), C.c_double(spacing),
                                                  _ptr(out, C.c_uint8)), "occ_voxel_parity")
        return out.reshape(nz, ny, nx)

    def cell_parity(self, cells, origin, spacing):
        """Inside attribute (1.0 / 0.0) of octree cells n x 4 (x, y, z, extent)."""
        cells = _i(cells).reshape(-1, 4)
        out = np.zeros(cells.shape[0], dtype=np.float32)
        self._schk(lib().hexopt_occ_cell_parity(C.c_void_p(self.h), cells.shape[0], _ptr(cells, C.c_int32),
                                                 _ptr(_d(origin)), C.c_double(spacing), _ptr(out, C.c_float)),
                   "occ_cell_parity")
        return out


class Curves:
    """Feature curves of the input surface: lists of surface vertex ids, closed when circle."""

    def __init__(self, V, curves, circles, device=0):
        self.V = _d(V)
        self.off = _i(np.concatenate([[0], np.cumsum([len(c) for c in curves])]))
        self.vs = _i(np.concatenate([np.asarray(c, dtype=np.int32) for c in curves]))
        self.circ = np.ascontiguousarray(circles, dtype=np.uint8)
        self.h = lib().hexopt_curves_create(self.V.shape[0], self.V.ctypes.data, len(curves), self.off.ctypes.data,
                                            self.vs.ctypes.data, self.circ.ctypes.data, device)
        if not self.h:
            raise RuntimeError("hexopt_curves_create: " + lib().hexopt_curves_error().decode())

    def close(self):
        if self.h:
            lib().hexopt_curves_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Octree:
    """An octree built on a Surface's device (hexopt_octree_build); refine() runs the targeted
    subdivision of listed cells (hexopt_octree_refine)."""

    def __init__(self, surface, grid, origin, voxel_size, mesh_transform=None, stop_extent=1, graded=True,
                 paired=True):
        g = np.ascontiguousarray(grid, dtype=np.int32).reshape(3)
        mt = None if mesh_transform is None else _d(mesh_transform)
        self.h = lib().hexopt_octree_build(C.c_void_p(surface.h), g.ctypes.data, _d(origin).ctypes.data,
                                           float(voxel_size), None if mt is None else mt.ctypes.data,
                                           int(stop_extent), int(bool(graded)), int(bool(paired)))
        if not self.h:
            raise RuntimeError("hexopt_octree_build: " + lib().hexopt_octree_error().decode())

    def refine(self, surface, stop_extent, cells):
        c = _i(cells).reshape(-1, 4)
        rc = lib().hexopt_octree_refine(C.c_void_p(self.h), C.c_void_p(surface.h), int(stop_extent), c.shape[0],
                                        c.ctypes.data)
        if rc:
            raise RuntimeError("hexopt_octree_refine (%d): %s" % (rc, lib().hexopt_octree_error().decode()))

    def result(self):
        sz = (C.c_int64 * 4)()
        ms = C.c_double()
        lib().hexopt_octree_sizes(C.c_void_p(self.h), sz, C.byref(ms))
        n, m = int(sz[0]), int(sz[1])
        leaves = np.empty((n, 4), np.int32)
        nodes = np.empty((m, 3), np.int32)
        V = np.empty((m, 3), np.float64)
        H = np.empty((n, 8), np.int32)
        lib().hexopt_octree_copy(C.c_void_p(self.h), leaves.ctypes.data, nodes.ctypes.data, V.ctypes.data,
                                 H.ctypes.data)
        return dict(leaves=leaves, nodes=nodes, V=V, H=H, levels=int(sz[2]), evaluated=int(sz[3]), ms=ms.value)

    def close(self):
        if self.h:
            lib().hexopt_octree_destroy(C.c_void_p(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
