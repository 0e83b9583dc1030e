"""Differentiable FMM potential: a torch.autograd.Function over the engine.

The potential of the FMM is phi_i = sum_{j != i} q_j / |x_i - x_j|.  For a
loss L(phi) with g = dL/dphi:

  dL/dq_j = sum_{i != j} g_i / |x_i - x_j| = psi_j, the potential of the
            charges g.  The 1/r operator is symmetric, so the adjoint FMM is
            the forward FMM with charges g on the same tree;
  dL/dx_j = g_j grad(phi)(x_j) + q_j grad(psi)(x_j): the first term from
            phi_j's own dependence on x_j, the second from phi_i (i != j)
            through source j.

The forward pass keeps only grad(phi) (3N values).  The backward pass sets the
charges g on the forward's tree (no rebuild) and recomputes every expansion
instead of storing them.  With an engine in HFMM_PRECISION_MIXED the adjoint
runs in the mixed form (FP32 P2P, split-bf16 M2L).

Tensors cross the boundary through host memory (CPU tensors, float64): the
engine owns its device memory and its own HIP runtime.  The reference has no
counterpart (its kernels have no derivative with respect to the charges or
the positions).
"""
from __future__ import annotations

import numpy as np
import torch


class FmmPotential(torch.autograd.Function):
    """phi = FmmPotential.apply(x, q, engine); x (N, 3) and q (N,) CPU tensors."""

    @staticmethod
    def forward(ctx, x, q, engine):
        xn = np.ascontiguousarray(x.detach().to(torch.float64).cpu().numpy())
        qn = np.ascontiguousarray(q.detach().to(torch.float64).cpu().numpy())
        engine.set_particles(xn[:, 0], xn[:, 1], xn[:, 2], qn)
        engine.build_tree()
        engine.evaluate()
        pot = engine.results()[0]
        ctx.engine = engine
        ctx.stamp = engine.tree_stamp
        ctx.xn, ctx.qn = xn, qn
        ctx.field = np.stack(engine.field(), axis=1)
        ctx.dtype = x.dtype
        return torch.from_numpy(pot).to(x.dtype)

    @staticmethod
    def backward(ctx, g):
        e = ctx.engine
        if e.tree_stamp != ctx.stamp:  # the engine ran another problem since the forward
            e.set_particles(ctx.xn[:, 0], ctx.xn[:, 1], ctx.xn[:, 2], ctx.qn)
            e.build_tree()
        gn = np.ascontiguousarray(g.detach().to(torch.float64).cpu().numpy())
        e.set_charges(gn)
        e.evaluate()
        psi = e.results()[0]
        field_g = np.stack(e.field(), axis=1)
        dx = gn[:, None] * ctx.field + ctx.qn[:, None] * field_g
        return torch.from_numpy(dx).to(ctx.dtype), torch.from_numpy(psi).to(ctx.dtype), None


def fmm_potential(x: torch.Tensor, q: torch.Tensor, engine) -> torch.Tensor:
    """Differentiable potential of charges q at positions x (one engine per problem size)."""
    return FmmPotential.apply(x, q, engine)
