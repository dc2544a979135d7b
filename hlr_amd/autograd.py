"""torch.autograd binding of the H-MVM and its adjoint (test binding).

`HMatVec.apply(x, plan, *arenas)` computes y = A x with A given by `plan`
whose stored factors are the tensors `arenas` (one per storage format that
the plan uses, in the plan's own layout, see arena_tensor below).  The
backward pass calls hmvm_grad: dL/dx = A^T dL/dy (transposed H-MVM) and
dL/d(factors) = per-block outer products with V^T x recomputed.  Gradients
with respect to factors are defined for FP64 and FP32 storage.
"""
from __future__ import annotations

import torch

from . import _native as N
from .plan import Plan

_DTYPES = (torch.float64, torch.float32, torch.uint16, torch.uint8)


def arena_tensor(plan: Plan, fmt: int) -> torch.Tensor:
    """Device copy of the stored arena of format fmt, as a torch tensor that
    can be handed to HMatVec as a leaf parameter."""
    import ctypes as C
    from .plan import _stream_handle
    n = plan.arena_elems(fmt)
    dt = torch.float64 if fmt == N.FMT_F64 else torch.float32
    if fmt not in (N.FMT_F64, N.FMT_F32):
        raise ValueError("factor gradients are defined for FP64 and FP32 storage only")
    t = torch.empty(n, dtype=dt, device=f"cuda:{plan.device}")
    if n:
        N.check(N.hmvm().hmvm_plan_get_arena(plan.handle, fmt, C.c_void_p(t.data_ptr()), _stream_handle(None)),
                "hmvm_plan_get_arena")
    return t


class HMatVec(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, plan: Plan, fmts, *arenas):
        for f, a in zip(fmts, arenas):
            plan.set_arena(f, a.contiguous())
        y = torch.zeros(plan.nrows, dtype=x.dtype, device=x.device)
        plan.apply(1.0, N.OP_NORMAL, x, y)
        ctx.plan = plan
        ctx.fmts = fmts
        ctx.save_for_backward(x, *arenas)
        return y

    @staticmethod
    def backward(ctx, gy):
        x, *arenas = ctx.saved_tensors
        plan = ctx.plan
        for f, a in zip(ctx.fmts, arenas):
            plan.set_arena(f, a.contiguous())
        gx = torch.zeros_like(x)
        want = any(ctx.needs_input_grad[3:])
        g = plan.grad(x, gy.contiguous().to(x.dtype), gx, factors=want)
        grads = []
        for f, a in zip(ctx.fmts, arenas):
            grads.append(g[f].to(a.dtype) if want and g[f] is not None else None)
        return (gx if ctx.needs_input_grad[0] else None), None, None, *grads


def hmatvec(x: torch.Tensor, plan: Plan, arenas: dict[int, torch.Tensor] | None = None) -> torch.Tensor:
    """y = A x, differentiable in x and (if given) in the factor arenas."""
    arenas = arenas or {}
    fmts = tuple(arenas.keys())
    return HMatVec.apply(x, plan, fmts, *[arenas[f] for f in fmts])
