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
charges g on the forward's tree (no rebuild), recomputes every expansion
instead of storing them, and puts the forward's charges back (the engine's
results() then hold the adjoint potential psi until its next evaluate()).
With an engine in HFMM_PRECISION_MIXED the adjoint runs in the mixed form
(FP32 P2P, split-bf16 M2L).

FmmPotential takes CPU tensors (float64 host copies cross the boundary);
fmm_potential() sends device tensors to the libtorch C++ autograd function
(csrc/torch_potential.cpp) instead, which reads them in place.  The reference
has no counterpart (its kernels have no derivative with respect to the
charges or the positions).
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
        # the forward's charges back on the tree: a later evaluate() computes
        # phi again (results() keep the adjoint potential psi until then)
        e.set_charges(ctx.qn)
        dx = gn[:, None] * ctx.field + ctx.qn[:, None] * field_g
        return torch.from_numpy(dx).to(ctx.dtype), torch.from_numpy(psi).to(ctx.dtype), None


def _device_function():
    """The libtorch C++ autograd function (scalfmm_amd/_lib/_torch_fmm): device
    tensors in and out, backward written in C++ (csrc/torch_potential.cpp)."""
    from ._lib import _torch_fmm
    if not _torch_fmm.same_runtime():
        raise RuntimeError("libscalfmm_hip.so was loaded before torch: the engine and torch use two HIP "
                           "runtimes and cannot share device tensors; import torch first")
    return _torch_fmm.potential


def fmm_potential(x: torch.Tensor, q: torch.Tensor, engine) -> torch.Tensor:
    """Differentiable potential of charges q at positions x (one engine per
    problem size).  Device tensors go through the C++ autograd function
    (the engine reads them in place, and after the backward it keeps the
    autograd context's copies of x and q as its particles: set them again
    before evaluating with the engine directly); CPU tensors through
    FmmPotential (host round trips)."""
    if x.is_cuda:
        return _device_function()(x, q, engine.h.value)
    return FmmPotential.apply(x, q, engine)
