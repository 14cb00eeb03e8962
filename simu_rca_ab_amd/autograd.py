"""Differentiable beamforming: the forward DAS kernel with the hand-written
adjoint kernel as its backward.

The forward is complex-linear in the I/Q data, y = A x, so the vector-Jacobian
product torch expects for complex tensors is grad_x = A^H grad_y; the geometry
is constant.  The adjoint kernel computes A^H by transposed interpolation
(scatter-add of q_k * conj(w e^{j 2 pi p}) * g into the traces).
"""
from __future__ import annotations

import torch

from .plan import RcaPlan


class DasFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, iq: torch.Tensor, plan: RcaPlan, compound: bool, deterministic: bool):
        out, _ = plan.forward(iq.contiguous(), compound=compound, want_mask=False)
        ctx.plan = plan
        ctx.deterministic = deterministic
        return out

    @staticmethod
    def backward(ctx, grad_out):
        grad = None
        if ctx.needs_input_grad[0]:
            grad = ctx.plan.adjoint(grad_out.contiguous().to(torch.complex64),
                                    deterministic=ctx.deterministic)
        return grad, None, None, None


def das(iq: torch.Tensor, plan: RcaPlan, compound: bool = True,
        deterministic: bool = False) -> torch.Tensor:
    """Beamform complex64 iq [F, T, C, Ns] -> [F, 1|T, P], differentiable in iq.
    deterministic: the backward uses the fixed-point adjoint (bitwise
    reproducible gradients, a workspace of 16 B per I/Q sample)."""
    return DasFunction.apply(iq, plan, compound, deterministic)
