"""Differentiable Joseph projector, exact adjoint and total variation.

The reference differentiates by hand: the least-squares gradient of
1/2||Ax-b||^2 is its adjoint backprojector applied to the residual
(emil_code/tomo3D.py:467-472 with view_proj3D.sinobackproject_linear,
view_proj3D.f:1096-1370) and the TV gradient is totalvar.total_var_grad3d
(totalvar.f:476-602, called from tomo3D.py:1034).  Here the same kernels sit
behind torch.autograd.Function so any torch optimiser can drive them:

* ``project(vol, op)``      y = A vol,        d/dvol <- A^T grad_y
* ``backproject(sino, op)`` v = A^T sino,     d/dsino <- A grad_v
* ``total_variation(vol, lam, eps)`` the value of totalvar.f:35-62
  (FP64 sum), gradient from totalvar.f:476-602 unnormalised.

``op`` is a ConeBeamOperator; A and A^T use the same matrix entries (the
adjoint is exact), so the gradients are exact up to FP32 rounding.
"""
from __future__ import annotations

import torch

from . import _dev
from ._lib import CT2_TV_WORK_DOUBLES, call


class _Project(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vol, op):
        ctx.op = op
        ctx.vol_dtype = vol.dtype
        return op.forward(vol.detach().contiguous())

    @staticmethod
    def backward(ctx, grad):
        if not ctx.needs_input_grad[0]:
            return None, None
        g = ctx.op.adjoint(grad.detach().to(torch.float32).contiguous())
        return g.to(ctx.vol_dtype), None


class _Backproject(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sino, op):
        ctx.op = op
        return op.adjoint(sino.detach().to(torch.float32).contiguous())

    @staticmethod
    def backward(ctx, grad):
        if not ctx.needs_input_grad[0]:
            return None, None
        return ctx.op.forward(grad.detach().to(torch.float32).contiguous()), None


def _tv_work(device):
    return torch.empty(CT2_TV_WORK_DOUBLES, dtype=torch.float64, device=device)


def tv_value_dev(vol, lam=0.0, eps=1e-9, out=None, work=None):
    """totalvar.f:35-62 on a device volume; returns a (1,) float64 tensor."""
    nx, ny, nz = vol.shape
    out = torch.empty((1,), dtype=torch.float64, device=vol.device) if out is None else out
    work = _tv_work(vol.device) if work is None else work
    call("ct2_tv_value", float(lam), _dev.ptr(vol), nx, ny, nz, float(eps), _dev.ptr(out),
         _dev.ptr(work), _dev.stream_handle())
    return out


def tv_grad_dev(vol, lam=0.0, eps=1e-9, normalizeq=0, out=None, bobo=None, work=None):
    """totalvar.f:476-602 on a device volume with a caller workspace; returns
    (gradient, bobo) without a host synchronisation."""
    nx, ny, nz = vol.shape
    g = torch.empty_like(vol) if out is None else out
    b = torch.empty((1,), dtype=torch.float32, device=vol.device) if bobo is None else bobo
    work = _tv_work(vol.device) if work is None else work
    call("ct2_tv_grad", float(lam), _dev.ptr(g), _dev.ptr(vol), nx, ny, nz, int(normalizeq),
         float(eps), _dev.ptr(b), _dev.ptr(work), _dev.stream_handle())
    return g, b


class _TotalVariation(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vol, lam, eps):
        v = vol.detach().to(torch.float32).contiguous()
        ctx.save_for_backward(v)
        ctx.lam, ctx.eps, ctx.vol_dtype = lam, eps, vol.dtype
        return tv_value_dev(v, lam, eps).to(torch.float32).reshape(())

    @staticmethod
    def backward(ctx, grad):
        if not ctx.needs_input_grad[0]:
            return None, None, None
        (v,) = ctx.saved_tensors
        g, _ = tv_grad_dev(v, ctx.lam, ctx.eps, normalizeq=0)
        return (g * grad.to(torch.float32)).to(ctx.vol_dtype), None, None


def project(vol, op):
    """A vol with gradient A^T (exact adjoint)."""
    return _Project.apply(vol, op)


def backproject(sino, op):
    """A^T sino with gradient A."""
    return _Backproject.apply(sino, op)


def total_variation(vol, lam=0.0, eps=1e-9):
    """TV(vol) of totalvar.f:35-62 (the regulariser whose gradient is
    totalvar.f:476-602), differentiable."""
    return _TotalVariation.apply(vol, float(lam), float(eps))


class Projector(torch.nn.Module):
    """nn.Module form of ``project`` for use inside models."""

    def __init__(self, op):
        super().__init__()
        self.op = op

    def forward(self, vol):
        return project(vol, self.op)
