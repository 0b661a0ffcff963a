"""Hand-written adjoint of the symmetric-Dirichlet SLIM energy: gradient and Hessian-vector
products (hexopt_energy_grad / hexopt_hvp) against central differences and the oracle's
SVD-form analytic gradient."""
import numpy as np
import pytest

import oracle_py as O
from helpers import constraint_kwargs, gpu_constraint_kwargs, load, rel

pytestmark = pytest.mark.gpu


def jittered(n, seed, invert_frac=0.0):
    from pyhexopt import meshgen as mg
    V, H = mg.grid_hex(n)
    bm = mg.grid_boundary_mask(n)
    V = mg.jitter(V, ~bm, 0.15 * 2.0 / n, seed=seed)
    if invert_frac:
        V, _ = mg.invert_hexes(V, H, invert_frac, seed=seed)
    return V, H


def setup(V, H, c=None):
    from pyhexopt import capi
    h = capi.HexOpt(V, H)
    h.reference()
    if c is not None:
        h.set_constraints(**gpu_constraint_kwargs(c))
    return h


@pytest.mark.parametrize("invert", [0.0, 0.02])
def test_gradient_vs_oracle(invert):
    V, H = jittered(8, 3, invert)
    h = setup(V, H)
    G = O.grad_ref(V, H)
    X = V + 0.003 * np.random.default_rng(5).normal(size=V.shape)
    E, go = O.sd_grad(H, G, X)
    g = h.energy_grad(X)
    assert rel(g, go) < 1e-10


def test_gradient_with_constraints_vs_oracle():
    c = load("slim_grid6_features")
    h = setup(c["V"], c["H"], c)
    G = O.grad_ref(c["V"], c["H"])
    rng = np.random.default_rng(1)
    X = c["V"] + 0.01 * rng.normal(size=c["V"].shape)
    aL = 0.01 * rng.normal(size=c["idsL"].size)
    kw = constraint_kwargs(c)
    E, go = O.sd_grad(c["H"], G, X, M=float(c["weight_opt"]), aL=aL, **kw)
    g = h.energy_grad(X, weight_opt=float(c["weight_opt"]), aL=aL)
    assert rel(g, go) < 1e-10


def test_gradient_central_differences():
    """Central differences (eps = 1e-6) of the GPU energy on 1,000 seeded coordinates."""
    V, H = jittered(8, 4, 0.02)
    h = setup(V, H)
    X = V.copy()
    g = h.energy_grad(X)
    rng = np.random.default_rng(0)
    picks = rng.choice(X.shape[0] * 3, size=1000, replace=False)
    eps = 1e-6
    errs = []
    for k in picks[:200]:   # each FD pair is two full energy launches; 200 keeps it quick
        v, d = divmod(int(k), 3)
        Xp = X.copy(); Xp[v, d] += eps
        Xm = X.copy(); Xm[v, d] -= eps
        h.set_positions(1, Xp); ep = sum(h.energy())
        h.set_positions(1, Xm); em = sum(h.energy())
        fd = (ep - em) / (2 * eps)
        errs.append(abs(fd - g[v, d]) / max(1e-3 * np.abs(g).max(), abs(g[v, d])))
    assert max(errs) < 1e-5


def test_hvp_vs_gradient_differences_and_symmetry():
    V, H = jittered(6, 7)
    h = setup(V, H)
    rng = np.random.default_rng(2)
    X = V.copy()
    u = rng.normal(size=X.shape)
    w = rng.normal(size=X.shape)
    Hu = h.hvp(X, u)
    Hw = h.hvp(X, w)
    eps = 1e-6
    fd = (h.energy_grad(X + eps * u) - h.energy_grad(X - eps * u)) / (2 * eps)
    assert rel(Hu, fd) < 1e-6
    a, b = (w * Hu).sum(), (u * Hw).sum()
    assert abs(a - b) <= 1e-10 * max(abs(a), abs(b))


def test_fp32_gradient_error_vs_fp64():
    V, H = jittered(8, 9)
    h = setup(V, H)
    g64 = h.energy_grad(V, precision=64)
    g32 = h.energy_grad(V.astype(np.float32), precision=32).astype(np.float64)
    # FP32 bulk arithmetic: relative error bound stated for the precision form
    assert rel(g32, g64) < 5e-4
