"""torch.autograd binding of the discrete adjoint (imports torch; not imported by the package).

    from hrweno_hip import Burgers1D, Adjoint
    from hrweno_hip.autograd import evolve
    adj = Adjoint(Burgers1D(n), nsteps=400, dt=dt, ckpt_every=20)
    uN = evolve(u0, adj)            # differentiable in u0 (float64 device tensor)
    loss = 0.5 * ((uN - target) ** 2).sum(); loss.backward()

The backward pass calls the hand-written reverse kernels through the C-ABI; if the adjoint
object has recorded another forward since (gradcheck does), the forward is recorded again
from the saved input first.
"""
from __future__ import annotations

import torch


class _Evolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, adj):
        uN = adj.forward(u0.detach().contiguous())
        ctx.adj = adj
        ctx.token = adj.token
        ctx.save_for_backward(u0)
        return uN

    @staticmethod
    def backward(ctx, gN):
        adj = ctx.adj
        if adj.token != ctx.token:
            (u0,) = ctx.saved_tensors
            adj.forward(u0.detach().contiguous())
            ctx.token = adj.token
        return adj.backward(gN.contiguous()), None


def evolve(u0: torch.Tensor, adj) -> torch.Tensor:
    """u_N = the adjoint object's nsteps forward from u0, differentiable in u0."""
    return _Evolve.apply(u0, adj)
