"""hexopt_scaled_jacobian_noalloc (the entry point that allocates nothing) against
hexopt_scaled_jacobian on the same meshes, through the C ABI: per-hex values and the minimum
bit-exact, the sum of per-hex minima to 1e-13 relative (it is summed in another hex order).
Two calls in a row, the second on a smaller mesh, each on a stream of its own; the device
buffers and streams come from the HIP runtime itself, through ctypes."""
import numpy as np
import pytest

from pyhexopt import capi, meshgen as mg

pytestmark = pytest.mark.gpu


def _mesh(n, seed, flip):
    V, H = mg.grid_hex(n)
    bm = mg.grid_boundary_mask(n)
    V = mg.jitter(V, ~bm, 0.2 * 2.0 / n, seed=seed)
    H = H.copy()
    rng = np.random.default_rng(seed)
    bad = rng.choice(H.shape[0], size=flip, replace=False)
    H[bad, 0], H[bad, 1] = H[bad, 1].copy(), H[bad, 0].copy()   # a few inverted hexes
    return V, H


class _Hip:
    """The HIP runtime libhexopt.so runs on, through ctypes: device buffers and a stream of
    the caller's own (the entry point's caller allocates its inputs and outputs)."""

    def __init__(self):
        import ctypes as C
        self.C = C
        capi.lib()       # libhexopt.so first: its RUNPATH names the runtime to share
        for name in ("libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so.7", "libamdhip64.so"):
            try:
                self.L = C.CDLL(name)
                break
            except OSError:
                continue
        else:
            raise RuntimeError("HIP runtime library not found")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.L.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.L.hipStreamDestroy.argtypes = [C.c_void_p]
        self.L.hipSetDevice.argtypes = [C.c_int]
        assert self.L.hipSetDevice(0) == 0

    def upload(self, a):
        C = self.C
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0      # host to device
        return p.value

    def download(self, ptr, out):
        assert self.L.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
        return out

    def free(self, *ptrs):
        for q in ptrs:
            self.L.hipFree(q)

    def stream(self):
        s = self.C.c_void_p()
        assert self.L.hipStreamCreate(self.C.byref(s)) == 0
        return s.value


def _noalloc(V, H, hip):
    X = hip.upload(np.ascontiguousarray(V.T))                              # x[d*nV + v]
    conn = hip.upload(np.ascontiguousarray(H.T.astype(np.int32)))          # conn[k*nH + h]
    per_hex = hip.upload(np.full(H.shape[0], -7.0))
    min_sum = hip.upload(np.full(2, -7.0))
    st = hip.stream()
    try:
        capi.scaled_jacobian_noalloc(0, V.shape[0], X, H.shape[0], conn, per_hex, min_sum, st)
        assert hip.L.hipStreamSynchronize(st) == 0
        return hip.download(per_hex, np.empty(H.shape[0])), hip.download(min_sum, np.empty(2))
    finally:
        hip.L.hipStreamDestroy(st)
        hip.free(X, conn, per_hex, min_sum)


def test_noalloc_scaled_jacobian_matches_context_entry_point():
    hip = _Hip()
    for n, seed, flip in ((14, 3, 40), (5, 4, 3)):          # second call: a smaller mesh
        V, H = _mesh(n, seed, flip)
        print("noalloc scaled Jacobian: grid %d^3, %d hexes" % (n, H.shape[0]), flush=True)
        ctx = capi.HexOpt(V, H, device=0)
        try:
            mn, avg, ph = ctx.scaled_jacobian(capi.CURRENT, per_hex=True)
        finally:
            ctx.close()
        ph2, ms = _noalloc(V, H, hip)
        assert mn < 0                                         # the flipped hexes are seen
        assert np.array_equal(ph2, ph)
        assert ms[0] == mn
        assert abs(ms[1] - avg * H.shape[0]) <= 1e-13 * abs(avg * H.shape[0])


def test_noalloc_rejects_oversized_and_null_input():
    with pytest.raises(RuntimeError, match="16,777,216"):
        capi.scaled_jacobian_noalloc(0, 8, 1 << 20, (1 << 24) + 1, 1 << 20, 0, 1 << 20)
    with pytest.raises(RuntimeError, match="null"):
        capi.scaled_jacobian_noalloc(0, 8, 0, 1, 1 << 20, 0, 1 << 20)
