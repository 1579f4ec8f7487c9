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


# ---- p-variants (totalvar.f:4-472) ----
def _work(device):
    from ._lib import CT2_TV_WORK_DOUBLES
    return torch.empty(CT2_TV_WORK_DOUBLES, dtype=torch.float64, device=device)


def gmi(smat, nx, ny, nz, eps=EPS):
    """totalvar.f:4-31; called from emil_code/tomo3D.py:1028."""
    host = _dev.is_host(smat)
    vol = _dev.to_dev(smat, torch.float32, (nx, ny, nz))
    out = torch.empty_like(vol)
    call("ct2_gmi", _dev.ptr(out), _dev.ptr(vol), nx, ny, nz, eps, _dev.stream_handle())
    return _dev.out(out, host)


def _value(sym, p, smat, nx, ny, nz, *extra, eps=EPS):
    vol = _dev.to_dev(smat, torch.float32, (nx, ny, nz))
    out = torch.empty((1,), dtype=torch.float64, device=vol.device)
    work = _work(vol.device)
    call(sym, p, _dev.ptr(vol), nx, ny, nz, *extra, eps, _dev.ptr(out), _dev.ptr(work),
         _dev.stream_handle())
    return float(out.item())


def calc_lp(p, smat, nx, ny, nz, eps=EPS):
    """totalvar.f:65-87; called from emil_code/tomo3D.py:971."""
    return _value("ct2_calc_lp", p, smat, nx, ny, nz, eps=eps)


def total_var3d_p(p, smat, nx, ny, nz, eps=EPS):
    """totalvar.f:91-117; called from emil_code/tomo3D.py:964."""
    return _value("ct2_total_var3d_p", p, smat, nx, ny, nz, eps=eps)


def total_dvar3d_p(p, smat, nx, ny, nz, dx, dy, dz, eps=EPS):
    """totalvar.f:121-146; called from emil_code/tomo3D.py:979."""
    return _value("ct2_total_dvar3d_p", p, smat, nx, ny, nz, dx, dy, dz, eps=eps)


def _grad(sym, p, smat, nx, ny, nz, *extra, normalizeq, eps=EPS):
    host = _dev.is_host(smat)
    vol = _dev.to_dev(smat, torch.float32, (nx, ny, nz))
    g = torch.empty_like(vol)
    bobo = torch.empty((1,), dtype=torch.float32, device=vol.device)
    work = _work(vol.device)
    call(sym, p, _dev.ptr(g), _dev.ptr(vol), nx, ny, nz, *extra, normalizeq, eps, _dev.ptr(bobo),
         _dev.ptr(work), _dev.stream_handle())
    return _dev.out(g, host), float(bobo.item())


def lp_grad3d(ppp, smat, nx, ny, nz, normalizeq, eps=EPS):
    """totalvar.f:150-215; called from emil_code/tomo3D.py:986.  Returns (tvgmat, bobo)."""
    return _grad("ct2_lp_grad3d", ppp, smat, nx, ny, nz, normalizeq=normalizeq, eps=eps)


def total_p_var_grad3d(ppp, smat, nx, ny, nz, normalizeq, eps=EPS):
    """totalvar.f:219-343; called from emil_code/tomo3D.py:994.  Returns (tvgmat, bobo)."""
    return _grad("ct2_total_p_var_grad3d", ppp, smat, nx, ny, nz, normalizeq=normalizeq,
                 eps=eps)


def total_p_dvar_grad3d(ppp, smat, nx, ny, nz, dx, dy, dz, normalizeq, eps=EPS):
    """totalvar.f:347-472; called from emil_code/tomo3D.py:1003.  Returns (tvgmat, bobo)."""
    return _grad("ct2_total_p_dvar_grad3d", ppp, smat, nx, ny, nz, dx, dy, dz,
                 normalizeq=normalizeq, eps=eps)
