"""+z ray-crossing occupancy at full size: the 256^3 voxel inside bitmap (16.8M voxels) of the
union of 8 seeded randomly rotated boxes (203,136 triangles, the soup of the 8 closed box
surfaces: compute_sign's signed crossing count makes a soup the union), bit-exact against
the CPU oracle's compute_sign(VoxelGrid) restatement (voxelization.h:221-272; fixture
tests/golden/occ_boxunion256.npz from tests/golden/make_boxunion.py)."""
import hashlib
import os

import numpy as np
import pytest

from helpers import load

pytestmark = pytest.mark.gpu


def test_voxel_parity_256_box_union():
    from pyhexopt import capi, meshgen as mg
    try:
        c = load("occ_boxunion256")
    except FileNotFoundError:
        pytest.skip("fixture not written yet (tests/golden/make_boxunion.py)")
    V, F, _ = mg.box_union_surface(8, 46, 0)
    h = hashlib.sha256(np.ascontiguousarray(V).tobytes() + np.ascontiguousarray(F).tobytes()).hexdigest()
    assert h == str(c["surface_sha256"])
    n = int(c["n"])
    s = capi.Surface(V, F)
    try:
        print("voxel parity %d^3 of %d triangles" % (n, F.shape[0]), flush=True)
        got = s.voxel_parity(n, n, n, c["origin"], float(c["spacing"]))
    finally:
        s.close()
    ref = np.unpackbits(c["bits"])[: n ** 3].reshape(n, n, n)
    diff = int((got != ref).sum())
    print("inside voxels %d (oracle %d), differing %d" % (int(got.sum()), int(c["count"]), diff), flush=True)
    assert diff == 0
