"""Differentiable sparse triangular solve: x = L^{-1} b.

TriangularSolve is a torch.autograd.Function over the C ABI of
libsympiler_amd.so.  Forward runs the GPU solve (symamd_solve); backward
runs the hand-written adjoint kernel (symamd_adjoint):

    grad_b      = L^{-T} grad_x        transposed solve in reverse level order
                                        (the reference's ltsolve,
                                        sparse_blas/sptrsv_lt.cpp:11-20)
    grad_L[p]   = -grad_b[i] * x[j]     for every stored entry p = (i, j) of L,
                                        written only on L's pattern, in the
                                        order of the values tensor

(from d x = -L^{-1} (dL) x).  FP64 and FP32 are supported; the plan's dtype
must match the values'.

One plan may serve several forwards with different values before their
backwards run: each forward gives the plan its values and notes the plan's
value generation; a backward that finds the generation moved on (another
forward, or any update_values, ran since) gives the plan the values its own
forward used (saved with the output) again.
"""
from __future__ import annotations

import torch

from . import Plan


class TriangularSolve(torch.autograd.Function):
    """x = TriangularSolve.apply(Lx, b, plan, executor)

    Lx: values of L on the plan's pattern (caller CSR order), 1-D tensor
    b:  right-hand side: a 1-D tensor of length n, or an (n, k) block whose
        columns are k right-hand sides (row-major); the backward pass of a
        block is the block adjoint, grad-L summed over the columns
    """

    @staticmethod
    def forward(ctx, Lx, b, plan: Plan, executor: str = "auto"):
        if Lx.dtype != plan.dtype or b.dtype != plan.dtype:
            raise ValueError("Lx and b must have the plan's dtype")
        Lx = Lx.detach().contiguous()
        plan.update_values(Lx)
        ctx.values_generation = plan.values_generation
        x = b.detach().contiguous().clone()
        if x.dim() == 1:
            plan.solve(x, executor=executor)
        else:
            plan.solve(x, layout="row", executor=executor)
        ctx.plan = plan
        ctx.executor = executor
        ctx.save_for_backward(x, Lx)
        return x

    @staticmethod
    def backward(ctx, gx):
        x, Lx = ctx.saved_tensors
        plan: Plan = ctx.plan
        if plan.values_generation != ctx.values_generation:
            plan.update_values(Lx)
        gx = gx.contiguous()
        gb = torch.empty_like(gx)
        gL = torch.empty(plan.nnz, dtype=gx.dtype, device=gx.device) \
            if ctx.needs_input_grad[0] else None
        if x.dim() == 1:
            plan.adjoint(x, gx, gb, gL, executor=ctx.executor)
        else:
            plan.adjoint_block(x, gx, gb, gL, layout="row", executor=ctx.executor)
        return gL, (gb if ctx.needs_input_grad[1] else None), None, None


def triangular_solve(Lx, b, plan: Plan, executor: str = "auto"):
    """Functional form of TriangularSolve."""
    return TriangularSolve.apply(Lx, b, plan, executor)
