"""Input/output of the compiled-reference driver (oracle/_ref/slim_ref).

Used to write the golden fixtures under tests/golden/ in the development container (where the
reference tree exists), and by bench.py's CPU-baseline leg, which times the built driver where
it travelled with the tree. GPU tests only read the committed fixtures.
"""
import os
import struct
import subprocess
import tempfile
import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF_BIN = os.path.join(REPO, "oracle", "_ref", "slim_ref")


def available():
    return os.path.exists(REF_BIN)


def _w(f, fmt, *v):
    f.write(struct.pack("<" + fmt, *v))


def _arr(f, a, dt):
    f.write(np.ascontiguousarray(a, dtype=dt).tobytes())


def read_out(path):
    out = {}
    with open(path, "rb") as f:
        while True:
            h = f.read(4)
            if not h:
                break
            (ln,) = struct.unpack("<I", h)
            name = f.read(ln).decode()
            kind = f.read(1).decode()
            (nb,) = struct.unpack("<Q", f.read(8))
            raw = f.read(nb)
            out[name] = np.frombuffer(raw, dtype=np.float64 if kind == "d" else np.int32).copy()
    return out


def write_slim_input(path, mode, V, H, energy=2, n_outer=1, n_inner=1, cube=True, projection=False, weight_opt=1.0,
                     lamT=0.0, lamT_bound=1e8, lamC=0.0, lamL=0.0, lamProj=1e5, hflag=None, b=None, bc=None,
                     idsT=None, normalT=None, disT=None, idsC=None, Cpos=None, idsL=None, axaL=None, originL=None,
                     surfV=None, surfF=None, tail=b""):
    """The driver's input file (slim_ref_driver.cpp's main, modes 0, 1, 3 and 7; `tail`: the
    bytes mode 7 reads after the surface)."""
    nV, nH = V.shape[0], H.shape[0]
    e3 = np.zeros((0, 3)); ei = np.zeros(0, dtype=np.int32)
    hflag = np.ones(nH, dtype=np.int32) if hflag is None else hflag
    with open(path, "wb") as f:
        _w(f, "i", mode)
        _w(f, "iiiiiii", nV, nH, energy, n_outer, n_inner, int(cube), int(projection))
        _w(f, "dddddd", weight_opt, lamT, lamT_bound, lamC, lamL, lamProj)
        _arr(f, V, "<f8"); _arr(f, H, "<i4"); _arr(f, hflag, "<i4")
        for ids, arrs in ((b, (bc,)), (idsT, (normalT, disT)), (idsC, (Cpos,)), (idsL, (axaL, originL))):
            ids = ei if ids is None else ids
            _w(f, "i", len(ids)); _arr(f, ids, "<i4")
            for a in arrs:
                if a is None:
                    a = e3[:0] if ids is None else np.zeros((len(ids), 3))
                _arr(f, a, "<f8")
        if surfV is None:
            _w(f, "ii", 0, 0)
        else:
            _w(f, "ii", surfV.shape[0], surfF.shape[0]); _arr(f, surfV, "<f8"); _arr(f, surfF, "<i4")
        f.write(tail)


def run_slim(mode, V, H, **kw):
    with tempfile.TemporaryDirectory() as td:
        pin, pout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        write_slim_input(pin, mode, V, H, **kw)
        subprocess.run([REF_BIN, pin, pout], check=True, stdout=subprocess.DEVNULL)
        return read_out(pout)


def run_feature(V, H, n_outer, weight_opt, lam0, lam_bound, idsT, idsC, Cpos, idsL, axaL, originL, surfV, surfF,
                fpatch, patchT, curves, circles, curveL):
    """Mode 7: the feature-alignment loop body with every vertex clean (lines re-projected
    onto their curves, planes onto their own patch, corners fixed, the three weights grown
    together from lam0)."""
    import io
    t = io.BytesIO()
    _arr(t, fpatch, "<i4"); _arr(t, patchT, "<i4")
    _w(t, "i", len(curves)); _arr(t, [len(c) for c in curves], "<i4")
    for c in curves:
        _arr(t, c, "<i4")
    _arr(t, circles, "<i4"); _arr(t, curveL, "<i4")
    nT = len(idsT)
    return run_slim(7, V, H, energy=2, n_outer=n_outer, n_inner=1, weight_opt=weight_opt, lamT=lam0,
                    lamT_bound=lam_bound, lamC=lam0, lamL=lam0, idsT=idsT, normalT=np.zeros((nT, 3)),
                    disT=np.zeros(nT), idsC=idsC, Cpos=Cpos, idsL=idsL, axaL=axaL, originL=originL, surfV=surfV,
                    surfF=surfF, tail=t.getvalue())


def run_glue(V, H, fixed, cubes, energy=1, weight_opt=1.0):
    """Mode 4: glueInternalNodes' exploded-mesh set-up and slim_solve2's gluing loop.
    cubes: (nH, 8, 3) exploded start positions (hex h, corner i)."""
    with tempfile.TemporaryDirectory() as td:
        pin, pout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(pin, "wb") as f:
            _w(f, "i", 4)
            _w(f, "iii", V.shape[0], H.shape[0], energy)
            _w(f, "d", weight_opt)
            _arr(f, V, "<f8"); _arr(f, H, "<i4"); _arr(f, fixed, "<i4"); _arr(f, cubes, "<f8")
        subprocess.run([REF_BIN, pin, pout], check=True, stdout=subprocess.DEVNULL)
        return read_out(pout)


def run_smooth(V, H, outer=3, inner=5, weight_opt=1.0):
    """Mode 5: the padding smoothing through the igl::slim variant (slim_opt_igl)."""
    with tempfile.TemporaryDirectory() as td:
        pin, pout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(pin, "wb") as f:
            _w(f, "i", 5)
            _w(f, "iiii", V.shape[0], H.shape[0], outer, inner)
            _w(f, "d", weight_opt)
            _arr(f, V, "<f8"); _arr(f, H, "<i4")
        subprocess.run([REF_BIN, pin, pout], check=True, stdout=subprocess.DEVNULL)
        return read_out(pout)


def run_aabb(V, F, P):
    with tempfile.TemporaryDirectory() as td:
        pin, pout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(pin, "wb") as f:
            _w(f, "i", 2)
            _w(f, "iii", V.shape[0], F.shape[0], P.shape[0])
            _arr(f, V, "<f8"); _arr(f, F, "<i4"); _arr(f, P, "<f8")
        subprocess.run([REF_BIN, pin, pout], check=True, stdout=subprocess.DEVNULL)
        return read_out(pout)


def run_point_line(A, B, P):
    """The reference's point_line_projection on (segment A-B, point P) triples (driver mode 6)."""
    with tempfile.TemporaryDirectory() as td:
        pin, pout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(pin, "wb") as f:
            _w(f, "i", 6)
            _w(f, "i", P.shape[0])
            _arr(f, A, "<f8"); _arr(f, B, "<f8"); _arr(f, P, "<f8")
        subprocess.run([REF_BIN, pin, pout], check=True, stdout=subprocess.DEVNULL)
        return read_out(pout)


METRO_BIN = os.path.join(REPO, "oracle", "_ref", "metro_ref")


def run_metro(V0, F0, V1, F1):
    """The reference's compute(mesh0, mesh1, bbox_diagonal, max, mean) (metro_hausdorff.cpp)."""
    with tempfile.TemporaryDirectory() as td:
        pin, pout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(pin, "wb") as f:
            for V, F in ((V0, F0), (V1, F1)):
                _w(f, "ii", V.shape[0], F.shape[0])
                _arr(f, V, "<f8")
                _arr(f, F, "<i4")
        subprocess.run([METRO_BIN, pin, pout], check=True, stdout=subprocess.DEVNULL)
        return read_out(pout)
