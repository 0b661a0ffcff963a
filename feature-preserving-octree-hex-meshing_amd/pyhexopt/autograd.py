"""torch.autograd wrapper of the hand-written adjoint (hexopt_energy_grad / hexopt_hvp).

`slim_energy(X, h)` returns the symmetric-Dirichlet energy plus the soft-constraint energy
of the HexOpt context `h` at positions X (nV x 3, float64). Its backward is the adjoint
gradient; the gradient's own backward is the Hessian-vector product, so
`torch.autograd.grad(..., create_graph=True)` gives Hv through a second backward. Values
travel through host memory: this is a harness for checking and driving the adjoint, not a
device-resident path.
"""
import torch

from . import capi


def _np(t):
    return t.detach().to("cpu", torch.float64).contiguous().numpy()


class _Gradient(torch.autograd.Function):
    @staticmethod
    def forward(ctx, X, h, weight_opt):
        ctx.h, ctx.weight_opt = h, weight_opt
        ctx.save_for_backward(X)
        return torch.from_numpy(h.energy_grad(_np(X), 64, weight_opt)).to(X)

    @staticmethod
    def backward(ctx, v):
        (X,) = ctx.saved_tensors
        hv = ctx.h.hvp(_np(X), _np(v), 64, ctx.weight_opt)
        return torch.from_numpy(hv).to(X), None, None


class SlimEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, X, h, weight_opt):
        ctx.h, ctx.weight_opt = h, weight_opt
        ctx.save_for_backward(X)
        h.set_positions(capi.CURRENT, _np(X))
        Eq, Es = h.energy(capi.SYMMETRIC_DIRICHLET, weight_opt)
        return X.new_tensor(Eq + Es)

    @staticmethod
    def backward(ctx, gE):
        (X,) = ctx.saved_tensors
        return gE * _Gradient.apply(X, ctx.h, ctx.weight_opt), None, None


def slim_energy(X, h, weight_opt=1.0):
    return SlimEnergy.apply(X, h, weight_opt)
