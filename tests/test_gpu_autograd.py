"""The adjoint behind torch.autograd: first and second backward of pyhexopt.autograd give
the adjoint gradient and Hessian-vector product, and torch's own finite-difference checks
(gradcheck, gradgradcheck) accept them on a small jittered grid with 2% inverted hexes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(n, seed):
    from pyhexopt import capi, meshgen as mg
    V, H = mg.grid_hex(n)
    V = mg.jitter(V, ~mg.grid_boundary_mask(n), 0.15 * 2.0 / n, seed=seed)
    V, _ = mg.invert_hexes(V, H, 0.02, seed=seed)
    h = capi.HexOpt(V, H)
    h.reference()
    return V, h


def test_backward_is_adjoint():
    import torch
    from pyhexopt.autograd import slim_energy
    V, h = _setup(6, 1)
    X = torch.tensor(V + 0.002, requires_grad=True)
    E = slim_energy(X, h)
    (g,) = torch.autograd.grad(E, X, create_graph=True)
    assert np.array_equal(g.detach().numpy(), h.energy_grad(V + 0.002))
    v = torch.tensor(np.random.default_rng(0).normal(size=V.shape))
    (hv,) = torch.autograd.grad((g * v).sum(), X)
    assert np.array_equal(hv.numpy(), h.hvp(V + 0.002, v.numpy()))


def test_gradcheck_and_gradgradcheck():
    import torch
    from pyhexopt.autograd import slim_energy
    V, h = _setup(3, 2)
    X = torch.tensor(V, requires_grad=True)
    f = lambda x: slim_energy(x, h)   # noqa: E731
    assert torch.autograd.gradcheck(f, (X,), eps=1e-6, atol=1e-4, rtol=1e-4)
    assert torch.autograd.gradgradcheck(f, (X,), eps=1e-6, atol=1e-4, rtol=1e-4)
