"""Differentiable NUFFTs: torch.autograd.Functions over the HIP library.

Type 2, c_j = sum_k f_k exp(i s k . x_j):
  grad_f = A^H g                  (a type-1 transform with -s, on a type-1
                                   plan of its own for NufftType2)
  grad_x = Re(conj(g) * dc/dx)    (dc/dx from the derivative-kernel interpolation)
Type 1, f_k = sum_j c_j exp(i s k . x_j):
  grad_c = A^H g                  (the plan's adjoint, a type-2 transform with -s)
  grad_x = Re(c * conj(d(A^H g)/dx))

g is PyTorch's gradient of a real loss with respect to a complex output
(dL/dRe + i dL/dIm). The reference has no gradient code; its adjoint
semantics are those of the CPU library's execute_adjoint
(include/finufft/execute.hpp:318-357).

Points are set on the plan for every forward call (they may change between
iterations); the backward pass reuses that plan and its sort, and sets the
forward call's points again when another call has replaced them since (two
forward calls of one module before a backward).
"""

from __future__ import annotations

import itertools

import torch

from ._plan import Plan

_FORWARD_IDS = itertools.count()


def _sum_batch(t, nbatch):
    return t.sum(dim=0) if nbatch else t


def _points_of(ctx):
    """The plan with the points of ctx's forward call set on it, and the
    forward's data tensor."""
    data, *pts = ctx.saved_tensors
    plan = ctx.plan
    if plan._pts_gen != ctx.pts_gen:
        plan.setpts(*[p.detach() for p in pts])
        ctx.pts_gen = plan._pts_gen
    return plan, data


class _AdjointSpreader:
    """The adjoint of a type-2 plan applied by a type-1 plan of its own
    (isign flipped; the same modes, vectors, tolerance and options): on the
    same points its spreading runs the type-1 kernels and bins (the column
    spreader for 3-D fp32), which stay fast on very dense point sets (the
    centre of radial spokes) where spreading on the type-2 plan's
    interpolation bins contends on LDS atomics. The plan is made at the
    first backward that needs it, and its points are set from the forward
    call's whenever they are not its last ones."""

    def __init__(self, n_modes, n_trans, eps, isign, dtype, opts):
        self._args = (tuple(n_modes), n_trans, eps, isign, dtype, dict(opts))
        self.plan = None
        self.key = None

    def __call__(self, g, pts, key):
        if self.plan is None:
            n_modes, n_trans, eps, isign, dtype, opts = self._args
            self.plan = Plan(1, n_modes, n_trans=n_trans, eps=eps, isign=-isign, dtype=dtype,
                             **opts)
        if self.key != key:
            self.plan.setpts(*[p.detach() for p in pts])
            self.key = key
        return self.plan.execute(g)


class _Type2(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f, plan, *pts):
        c = plan.execute(f)
        ctx.plan = plan
        ctx.pts_gen = plan._pts_gen
        ctx.fwd_key = next(_FORWARD_IDS)
        ctx.save_for_backward(f, *pts)
        return c

    @staticmethod
    def backward(ctx, g):
        plan, f = _points_of(ctx)
        g = g.contiguous()
        grad_f = None
        if ctx.needs_input_grad[0]:
            adj = getattr(plan, "adjoint_spreader", None)
            if adj is not None:
                grad_f = adj(g, ctx.saved_tensors[1:], ctx.fwd_key)
            else:
                grad_f = plan.execute_adjoint(g)
        grads = [None] * (len(ctx.needs_input_grad) - 2)
        if any(ctx.needs_input_grad[2:]):
            _, dc = plan.interp_grad(f, value=False)
            nb = g.dim() - 1
            for k in range(len(grads)):
                if ctx.needs_input_grad[2 + k]:
                    grads[k] = _sum_batch((g.conj() * dc[k]).real, nb)
        return (grad_f, None, *grads)


class _Type1(torch.autograd.Function):
    @staticmethod
    def forward(ctx, c, plan, *pts):
        f = plan.execute(c)
        ctx.plan = plan
        ctx.pts_gen = plan._pts_gen
        ctx.save_for_backward(c, *pts)
        return f

    @staticmethod
    def backward(ctx, g):
        plan, c = _points_of(ctx)
        g = g.contiguous()
        need_x = any(ctx.needs_input_grad[2:])
        grads = [None] * (len(ctx.needs_input_grad) - 2)
        if need_x:
            ag, dag = plan.interp_grad(g, value=ctx.needs_input_grad[0])
            nb = c.dim() - 1
            for k in range(len(grads)):
                if ctx.needs_input_grad[2 + k]:
                    grads[k] = _sum_batch((c * dag[k].conj()).real, nb)
            grad_c = ag
        else:
            grad_c = plan.execute_adjoint(g) if ctx.needs_input_grad[0] else None
        return (grad_c, None, *grads)


class NufftType2(torch.nn.Module):
    """c = type-2 NUFFT of modes f at points (x[, y[, z]]), differentiable in f
    and in the points. n_modes in array order; isign -1 by default (as the
    reference's type 2)."""

    def __init__(self, n_modes, eps=1e-6, isign=-1, n_trans=1, dtype="complex64", **opts):
        super().__init__()
        self.plan = Plan(2, tuple(n_modes), n_trans=n_trans, eps=eps, isign=isign, dtype=dtype,
                         **opts)
        self.plan.adjoint_spreader = _AdjointSpreader(n_modes, n_trans, eps, isign, dtype, opts)

    def forward(self, f, *pts):
        self.plan.setpts(*[p.detach() for p in pts])
        return _Type2.apply(f, self.plan, *pts)


class NufftType1(torch.nn.Module):
    """f = type-1 NUFFT of strengths c at points (x[, y[, z]]), differentiable
    in c and in the points. isign +1 by default (as the reference's type 1)."""

    def __init__(self, n_modes, eps=1e-6, isign=1, n_trans=1, dtype="complex64", **opts):
        super().__init__()
        self.plan = Plan(1, tuple(n_modes), n_trans=n_trans, eps=eps, isign=isign, dtype=dtype,
                         **opts)

    def forward(self, c, *pts):
        self.plan.setpts(*[p.detach() for p in pts])
        return _Type1.apply(c, self.plan, *pts)


def nufft2(f, *pts, eps=1e-6, isign=-1, **opts):
    """One-shot differentiable type 2: f has shape n_modes or (n_trans, *n_modes)."""
    dim = len(pts)
    n_modes = tuple(f.shape[-dim:])
    n_trans = f.shape[0] if f.dim() > dim else 1
    dtype = "complex128" if f.dtype == torch.complex128 else "complex64"
    return NufftType2(n_modes, eps=eps, isign=isign, n_trans=n_trans, dtype=dtype, **opts)(f, *pts)


def nufft1(c, *pts, n_modes, eps=1e-6, isign=1, **opts):
    """One-shot differentiable type 1: c has shape (M,) or (n_trans, M)."""
    n_trans = c.shape[0] if c.dim() > 1 else 1
    dtype = "complex128" if c.dtype == torch.complex128 else "complex64"
    return NufftType1(tuple(n_modes), eps=eps, isign=isign, n_trans=n_trans, dtype=dtype,
                      **opts)(c, *pts)
