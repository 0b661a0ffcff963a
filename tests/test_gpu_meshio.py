"""The file formats in the GPU suite (row f4 of the scope table with a record on the GPU box):
the reference-byte checks of tests/test_meshio.py run again here, and a mesh optimised on the
device is written as VTK and Medit .mesh by the library's writers and read back: ids exactly,
every coordinate as float("%g" % x) (the six digits of the reference's stream output)."""
import numpy as np
import pytest

import test_meshio as tm
from helpers import gpu_constraint_kwargs, load
from pyhexopt import meshio as mio

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", tm._names("w"))
def test_writer_bytes_match_reference_on_gpu_box(name, tmp_path):
    tm.test_writer_bytes_match_reference(name, tmp_path)


@pytest.mark.parametrize("name", tm._names("r"))
def test_reader_matches_reference_on_gpu_box(name, tmp_path):
    tm.test_reader_matches_reference(name, tmp_path)


@pytest.mark.parametrize("fmt", ["vtk", "mesh"])
def test_write_device_optimised_mesh(fmt, tmp_path):
    from pyhexopt import capi
    c = load("slim_grid6_features")
    h = capi.HexOpt(c["V"], c["H"])
    try:
        h.reference()
        h.set_constraints(**gpu_constraint_kwargs(c))
        h.slim_solve(int(c["iters"]), weight_opt=float(c["weight_opt"]))
        X = h.get_positions()
    finally:
        h.close()
    p = str(tmp_path / ("opt." + fmt))
    if fmt == "vtk":
        mio.write_vtk(p, X, c["H"], mio.HEX)
        m = mio.read_vtk(p)
    else:
        mio.write_mesh(p, X, c["H"], mio.HEX)
        m = mio.read_mesh(p)
    want = np.array([float("%g" % x) for x in X.reshape(-1)]).reshape(X.shape)
    assert np.array_equal(m.V, want)
    assert np.array_equal(m.idx.reshape(-1, 8), c["H"])
