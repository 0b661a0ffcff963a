"""Area-uniform surface samples (meshgen.surface_samples), used by the sampled Hausdorff
measure of the scale workload. CPU only."""
import numpy as np


def test_samples_are_area_uniform():
    from pyhexopt import meshgen as mg
    V, F = mg.box_surface(2)
    P = mg.surface_samples(V, F, 60000, seed=1)
    # every face of the unit cube gets a sixth of the samples
    face = np.argmax(np.abs(P), axis=1) * 2 + (np.take_along_axis(P, np.argmax(np.abs(P), axis=1)[:, None], 1)[:, 0] > 0)
    counts = np.bincount(face, minlength=6)
    assert np.all(np.abs(counts / 60000 - 1 / 6) < 0.01)
    assert np.allclose(np.abs(P).max(axis=1), 0.5)
