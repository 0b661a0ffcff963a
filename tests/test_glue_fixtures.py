"""CPU checks of the glue fixtures: the synthetic start copies are reproducible from the
generator, and the reference run satisfies the loop's own rules (weight schedule and stop)."""
import numpy as np

from helpers import load


def test_exploded_start_reproducible():
    from pyhexopt import meshgen as mg
    for name, n, amax, seed in (("glue_grid4_untangle", 4, 0.2, 5), ("glue_grid5_tangled", 5, 0.3, 6)):
        c = load(name)
        V, H = mg.grid_hex(n)
        start = mg.exploded_start(V, H, c["fixed"].astype(bool), c["V"], amax, seed=seed)
        assert np.array_equal(start, c["start"])
        Ve, He, orig = mg.exploded_mesh(c["H"], c["start"])
        assert np.array_equal(Ve[He[:, 3]], c["start"][:, 3]) and np.array_equal(orig.reshape(-1, 8), c["H"])


def test_reference_weight_schedule():
    """lamda_glue starts at 1e4 and follows min(1e4 max(Eq/Es lam, 1), 1e16); the loop stops at
    the first positive minimum scaled Jacobian or after 50 iterations."""
    for name in ("glue_grid4_untangle", "glue_grid5_tangled"):
        log = load(name)["ref_log"].reshape(-1, 5)
        assert log[0, 3] == 1e4
        for k in range(1, log.shape[0]):
            lam = min(1e4 * max(log[k - 1, 1] / log[k - 1, 2] * log[k - 1, 3], 1.0), 1e16)
            assert log[k, 3] == lam
        assert (log[:-1, 4] <= 0).all() and (log[-1, 4] > 0 or log.shape[0] == 50)
