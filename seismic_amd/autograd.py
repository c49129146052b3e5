"""The propagator as a differentiable torch operation.

``propagate(vel, plan)`` maps a batch of parameter vectors (B, nz*nx+1) to
seismograms (B, shots*nt*receivers) on the GPU.  Its backward is the
hand-written adjoint pass: the forward keeps the wavefield every K steps
(plan ``ckpt_every=K``; 0 keeps the full history) and the backward recomputes
each segment from its checkpoint before running the reverse steps
(``Propagator.forward_save`` / ``Propagator.adjoint``).  If another pass ran
on the plan between forward and backward, the backward recomputes the forward
instead.  The gradient is J^T g for the incoming seismogram gradient g, the
same quantity as vel_to_seis(vec, vec_adjoint=g) in the reference
(core/seis_forward2.py:146-331), including d/d min_vel.
"""
from __future__ import annotations

import torch

from .propagator import Propagator


class _Propagate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vel: torch.Tensor, plan: Propagator):
        flat = vel.reshape(-1)
        seis, token = plan.forward_save(flat)
        ctx.plan = plan
        ctx.token = token
        ctx.save_for_backward(vel)
        return seis.reshape(vel.shape[0], -1) if vel.dim() == 2 else seis

    @staticmethod
    def backward(ctx, gseis: torch.Tensor):
        (vel,) = ctx.saved_tensors
        plan = ctx.plan
        flat = vel.reshape(-1)
        g = gseis.reshape(-1)
        if plan.saved_valid(ctx.token):
            _, grad = plan.adjoint(flat, g, ctx.token, on_residual=False)
        else:
            _, _, grad = plan.gradient(flat, g, on_residual=False)
        return grad.to(vel.dtype).reshape(vel.shape), None


def propagate(vel: torch.Tensor, plan: Propagator) -> torch.Tensor:
    """Seismograms of the plan's models; differentiable in ``vel``."""
    if vel.dim() == 2 and vel.shape[0] != plan.n_models:
        raise ValueError(f"batch of {vel.shape[0]} models, plan holds {plan.n_models}")
    return _Propagate.apply(vel, plan)


class SmallCNNPrior(torch.nn.Module):
    """A small convolutional prior on the velocity field: mean squared output
    of a 3-layer CNN on (v - 3000)/1000.  Stands in for a learned prior term
    in the loss; randomly initialised (seeded)."""

    def __init__(self, nz: int, nx: int, width: int = 8, seed: int = 0):
        super().__init__()
        gen = torch.Generator().manual_seed(seed)
        self.nz, self.nx = nz, nx
        self.net = torch.nn.Sequential(
            torch.nn.Conv2d(1, width, 3, padding=1), torch.nn.Tanh(),
            torch.nn.Conv2d(width, width, 3, padding=1), torch.nn.Tanh(),
            torch.nn.Conv2d(width, 1, 3, padding=1))
        with torch.no_grad():
            for p in self.net.parameters():
                p.copy_(torch.randn(p.shape, generator=gen) * 0.3)

    def forward(self, vel: torch.Tensor) -> torch.Tensor:
        """vel: (B, nz*nx+1) -> per-model prior value (B,)."""
        v = vel[:, :-1].reshape(-1, 1, self.nz, self.nx)
        return self.net((v - 3000.0) / 1000.0).pow(2).mean(dim=(1, 2, 3))
