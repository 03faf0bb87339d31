"""Differentiable projector and objectives on the HIP kernels.

* :class:`ForwardProjection` -- ``y = A_k x`` as a ``torch.autograd.Function`` whose backward is the
  hand-written backprojection kernel (``A_k^T``), so gradients of any torch objective of the
  projections flow through the library's own BP.
* :class:`RDPPrior` -- the relative difference prior value R(u) (omg_rdp_value) with the gradient
  kernel (omg_rdp_gradient, epsilon 0) as its backward.
* :func:`poisson_loglik` / :func:`penalised_nll` -- the Poisson log-likelihood
  ``L(x) = sum y log(Ax) - Ax`` and the objective ``-L(x) + R(x)`` whose OSL-OSEM iterations the
  reference runs (source/cpp/reconstructionAF.h; RDP.h:4-78). The reference never evaluates these
  values; they are defined here (RDP value: interior neighbour pairs once, pairs with a neighbour
  outside the volume (value 0) with weight 1, so that its gradient is the eps -> 0 limit of the
  reference's RDP gradient with zero boundary).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib


def _p(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class ForwardProjection(torch.autograd.Function):
    """Forward projection with the backprojection kernel as its adjoint."""

    @staticmethod
    def forward(ctx, x, proj, subset):
        ctx.proj = proj
        ctx.subset = subset
        return proj.forwardProject(x.detach().reshape(-1).contiguous(), subset)

    @staticmethod
    def backward(ctx, gy):
        gx = ctx.proj.backwardProject(gy.detach().reshape(-1).contiguous(), ctx.subset)
        return gx, None, None


def forward_project(x, proj, subset=-1):
    """Differentiable ``A_subset x`` (all planes when the projector has one subset)."""
    return ForwardProjection.apply(x, proj, subset)


class RDPPrior(torch.autograd.Function):
    """R(u) as a float64 scalar; backward uses the RDP gradient kernel with epsilon 0."""

    @staticmethod
    def forward(ctx, u, shape, gamma, beta):
        u = u.detach().reshape(-1).contiguous()
        out = torch.empty(1, dtype=torch.float64, device=u.device)
        _lib.call("omg_rdp_value", _p(u), *shape, float(gamma), float(beta), _p(out), _stream())
        ctx.save_for_backward(u)
        ctx.shape, ctx.gamma, ctx.beta = shape, gamma, beta
        return out[0]

    @staticmethod
    def backward(ctx, g):
        (u,) = ctx.saved_tensors
        grad = torch.zeros_like(u)
        _lib.call("omg_rdp_gradient", _p(u), _p(grad), *ctx.shape, float(ctx.gamma), 0.0, float(ctx.beta), 0,
                  _stream())
        return grad * g.to(grad.dtype), None, None, None


def rdp(u, shape, gamma=2.0, beta=0.1):
    return RDPPrior.apply(u, tuple(int(s) for s in shape), gamma, beta)


def poisson_loglik(x, proj, meas):
    """L(x) = sum_i y_i log((A x)_i) - (A x)_i over all subsets, accumulated in float64.

    ``meas`` is the full measurement vector in subset order. Differentiable in ``x``.
    """
    total = None
    for k in range(proj.subsets):
        yhat = forward_project(x, proj, k).double()
        off, n = int(proj.nTotMeas[k]), int(proj.nMeasSubset[k, 0])
        y = meas[off:off + n].double()
        term = torch.where(y != 0, y * torch.log(yhat), torch.zeros_like(yhat)).sum() - yhat.sum()
        total = term if total is None else total + term
    return total


def penalised_nll(x, proj, meas, gamma=2.0, beta=0.1):
    """-L(x) + R(x): the objective of OSL-OSEM with the RDP prior."""
    shape = (int(proj.Nx[0]), int(proj.Ny[0]), int(proj.Nz[0]))
    return -poisson_loglik(x, proj, meas) + rdp(x, shape, gamma, beta)
