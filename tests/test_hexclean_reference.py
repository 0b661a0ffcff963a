"""clean_hex_mesh's flag passes (include/hexclean.h, host code) against the reference's own
tagging_uneven_element, clean_non_manifold_ve and drop_small_pieces compiled from its
sources (tests/golden/clean_blobs.npz, written by tests/golden/make_clean.py): 40 seeded
random blobs of 6^3 to 14^3 cells, half renumbered at random, flag for flag after each pass,
then the medial-surface tags of the final flags (vertex flags, and the tagged faces with
their corners in the reference's face order). CPU only."""
import numpy as np
import pytest

from helpers import load
from pyhexopt import hexclean

C = load("clean_blobs")


@pytest.mark.parametrize("i", range(int(C["count"])))
def test_cleaning_passes_match_compiled_reference(i):
    nV, H, flag = int(C["m%d_nV" % i]), C["m%d_H" % i], C["m%d_in" % i]
    tag = hexclean.tagging_uneven_element(nV, H, flag)
    assert np.array_equal(tag, C["m%d_tag" % i])
    man = hexclean.clean_non_manifold_ve(nV, H, tag)
    assert np.array_equal(man, C["m%d_manifold" % i])
    fin = hexclean.drop_small_pieces(nV, H, man)
    assert np.array_equal(fin, C["m%d_final" % i])
    assert np.array_equal(hexclean.clean_flags(nV, H, flag), C["m%d_final" % i])
    vmed, fmed = hexclean.medial_surface(nV, H, fin)
    assert np.array_equal(vmed, C["m%d_medial_v" % i])
    assert np.array_equal(fmed, C["m%d_medial_f" % i])
