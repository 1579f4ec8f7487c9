"""Drop-in for the reference's f2py module ``view_proj3D_tomo``
(emil_code/view_proj3D_tomo.f, imported as ``vpt`` at emil_code/tomo3D.py:12):
the Joseph operators with the driving axis forced to z."""
from __future__ import annotations

import torch

from . import _dev
from ._lib import call
from .view_proj3D import _art, _em, _fwd, _inout


def sinoprojz(*a):
    """view_proj3D_tomo.f:4-148; dict key 'zproj_linear' (emil_code/tomo3D.py:66)."""
    return _fwd("ct2_sinoprojz", *a)


def artstepz(*a):
    """view_proj3D_tomo.f:154-347; dict key 'zart_linear'."""
    return _art("ct2_artstepz", *a)


def emstep_linearz(*a):
    """view_proj3D_tomo.f:615-822; dict key 'zem_linear'."""
    return _em("ct2_emstep_linearz", *a)


def artstepz2vol(reg_fac, sinomat, indsino, frame_vectors, ns, nu, nv, du, dv, u0, v0,
                 smat, dx, dy, dz, x0, y0, z0, nx, ny, nz,
                 smata, dxa, dya, dza, x0a, y0a, z0a, nxa, nya, nza):
    """view_proj3D_tomo.f:352-612; returns (smat, smata) like
    emil_code/tomo3D.py:1318."""
    host = _dev.is_host(sinomat, indsino, frame_vectors, smat, smata)
    ind = _dev.ind_dev(indsino, (ns, nu, nv))
    fv = _dev.to_dev(frame_vectors, torch.float32, (ns, 18))
    sino = _dev.to_dev(sinomat, torch.float32, (ns, nu, nv))
    vol = _inout(smat, (nx, ny, nz), host)
    vola = _inout(smata, (nxa, nya, nza), host)
    call("ct2_artstepz2vol", reg_fac, _dev.ptr(sino), _dev.ptr(ind), _dev.ptr(fv),
         ns, nu, nv, du, dv, u0, v0, _dev.ptr(vol), dx, dy, dz, x0, y0, z0, nx, ny, nz,
         _dev.ptr(vola), dxa, dya, dza, x0a, y0a, z0a, nxa, nya, nza, _dev.stream_handle())
    return _dev.out(vol, host), _dev.out(vola, host)
