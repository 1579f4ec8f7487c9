"""Drop-in for the reference's f2py module ``totalvar`` (emil_code/totalvar.f,
imported as ``tv`` at emil_code/tomo3D.py:13).  ``eps`` defaults to the
reference's hard-coded 1e-9 and can be overridden by keyword."""
from __future__ import annotations

import torch

from . import _dev
from ._lib import call

EPS = 1e-9


def total_var3d(ambdala, smat, nx, ny, nz, eps=EPS):
    """totalvar.f:35-62; called from emil_code/tomo3D.py:956.  Returns a float
    (the reference's Real*4 function value)."""
    vol = _dev.to_dev(smat, torch.float32, (nx, ny, nz))
    out = torch.empty((1,), dtype=torch.float64, device=vol.device)
    call("ct2_total_var3d", ambdala, _dev.ptr(vol), nx, ny, nz, eps, _dev.ptr(out),
         _dev.stream_handle())
    return float(out.item())


def total_var_grad3d(ambdala, smat, nx, ny, nz, normalizeq, eps=EPS):
    """totalvar.f:476-602; called from emil_code/tomo3D.py:1034.  Returns
    (tvgmat, bobo) like the f2py wrapper; bobo is a Python float."""
    host = _dev.is_host(smat)
    vol = _dev.to_dev(smat, torch.float32, (nx, ny, nz))
    g = torch.empty_like(vol)
    bobo = torch.empty((1,), dtype=torch.float32, device=vol.device)
    call("ct2_total_var_grad3d", ambdala, _dev.ptr(g), _dev.ptr(vol), nx, ny, nz, normalizeq,
         eps, _dev.ptr(bobo), _dev.stream_handle())
    return _dev.out(g, host), float(bobo.item())


def total_var_grad3d_dev(ambdala, vol, normalizeq, eps=EPS, out=None, bobo=None):
    """Device-resident variant for iteration loops: no host synchronisation.
    Returns (gradient tensor, bobo tensor of shape (1,))."""
    nx, ny, nz = vol.shape
    g = torch.empty_like(vol) if out is None else out
    b = torch.empty((1,), dtype=torch.float32, device=vol.device) if bobo is None else bobo
    call("ct2_total_var_grad3d", ambdala, _dev.ptr(g), _dev.ptr(vol), nx, ny, nz, normalizeq,
         eps, _dev.ptr(b), _dev.stream_handle())
    return g, b
