"""Drop-in for the reference's f2py module ``ray_proj3D``
(emil_code/ray_proj3D.f, imported at emil_code/tomo3D.py:11): Joseph
operators over an explicit ray list ray_vectors[nr,6] = (source, bin)."""
from __future__ import annotations

import torch

from . import _dev
from ._lib import call
from .view_proj3D import _inout


def _fwd(sym, indsino, ray_vectors, nr, smat, dx, dy, dz, x0, y0, z0, nx, ny, nz):
    host = _dev.is_host(indsino, ray_vectors, smat)
    ind = _dev.ind_dev(indsino, (nr,))
    rv = _dev.to_dev(ray_vectors, torch.float32, (nr, 6))
    vol = _dev.to_dev(smat, torch.float32, (nx, ny, nz))
    sino = torch.empty((nr,), dtype=torch.float32, device=vol.device)
    call(sym, _dev.ptr(sino), _dev.ptr(ind), _dev.ptr(rv), nr, _dev.ptr(vol),
         dx, dy, dz, x0, y0, z0, nx, ny, nz, _dev.stream_handle())
    return _dev.out(sino, host)


def sinoproj_linear(*a):
    """ray_proj3D.f:514-671; called from emil_code/tomo3D.py:1191."""
    return _fwd("ct2_ray_sinoproj_linear", *a)


def sinoproj_nearest(*a):
    """ray_proj3D.f:677-804."""
    return _fwd("ct2_ray_sinoproj_nearest", *a)


def sinobackproject_nearest(sinomat, indsino, ray_vectors, nr, dx, dy, dz, x0, y0, z0,
                            nx, ny, nz):
    """ray_proj3D.f:809-980 (bounds 0 <= i < n-1, ray_proj3D.f:895);
    called from emil_code/tomo3D.py:641."""
    host = _dev.is_host(sinomat, indsino, ray_vectors)
    ind = _dev.ind_dev(indsino, (nr,))
    rv = _dev.to_dev(ray_vectors, torch.float32, (nr, 6))
    sino = _dev.to_dev(sinomat, torch.float32, (nr,))
    vol = torch.empty((nx, ny, nz), dtype=torch.float32, device=sino.device)
    call("ct2_ray_sinobackproject_nearest", _dev.ptr(sino), _dev.ptr(ind), _dev.ptr(rv), nr,
         _dev.ptr(vol), dx, dy, dz, x0, y0, z0, nx, ny, nz, _dev.stream_handle())
    return _dev.out(vol, host)


def _art(sym, reg_fac, sinomat, indsino, ray_vectors, nr, smat, dx, dy, dz, x0, y0, z0,
         nx, ny, nz):
    host = _dev.is_host(sinomat, indsino, ray_vectors, smat)
    ind = _dev.ind_dev(indsino, (nr,))
    rv = _dev.to_dev(ray_vectors, torch.float32, (nr, 6))
    sino = _dev.to_dev(sinomat, torch.float32, (nr,))
    vol = _inout(smat, (nx, ny, nz), host)
    call(sym, reg_fac, _dev.ptr(sino), _dev.ptr(ind), _dev.ptr(rv), nr, _dev.ptr(vol),
         dx, dy, dz, x0, y0, z0, nx, ny, nz, _dev.stream_handle())
    return _dev.out(vol, host)


def artstep_linear(*a):
    """ray_proj3D.f:987-1267; called from emil_code/tomo3D.py:1347."""
    return _art("ct2_ray_artstep_linear", *a)


def artstep_nearest(*a):
    """ray_proj3D.f:1272-1449 (bounds < n-1)."""
    return _art("ct2_ray_artstep_nearest", *a)


def emstep_nearest(sinomat, indsino, ray_vectors, nr, smat, dx, dy, dz, x0, y0, z0,
                   nx, ny, nz):
    """ray_proj3D.f:1454-1651 (bounds < n-1); called from emil_code/tomo3D.py:1495."""
    host = _dev.is_host(sinomat, indsino, ray_vectors, smat)
    ind = _dev.ind_dev(indsino, (nr,))
    rv = _dev.to_dev(ray_vectors, torch.float32, (nr, 6))
    sino = _dev.to_dev(sinomat, torch.float32, (nr,))
    vol = _inout(smat, (nx, ny, nz), host)
    call("ct2_ray_emstep_nearest", _dev.ptr(sino), _dev.ptr(ind), _dev.ptr(rv), nr,
         _dev.ptr(vol), dx, dy, dz, x0, y0, z0, nx, ny, nz, _dev.stream_handle())
    return _dev.out(vol, host)


# ---- analytic phantoms (ray_proj3D.f:175-512) ----
def gen_shape_proj(indsino, ray_vectors, nr, att, nsurf, *surf):
    """ray_proj3D.f:175-431; called from emil_code/tomo3D.py:2688."""
    from . import quadric
    host = _dev.is_host(indsino, ray_vectors)
    t = quadric.tables_from_f2py(att, nsurf, *surf)
    ind = _dev.ind_dev(indsino, (nr,))
    rv = _dev.to_dev(ray_vectors, torch.float32, (nr, 6))
    sino = torch.empty((nr,), dtype=torch.float32, device=rv.device)
    quadric.project_rays(t, sino, ind, rv, nr, 0)
    return _dev.out(sino, host)


def ellipsoid_proj(indsino, ray_vectors, nr, ax, ay, az, x0, y0, z0, alpha, beta, att):
    """ray_proj3D.f:435-512; called from emil_code/tomo3D.py:2891."""
    from . import quadric
    host = _dev.is_host(indsino, ray_vectors)
    t = quadric.Tables()
    t.add_ellipsoid(att, ax, ay, az, x0, y0, z0, alpha, beta)
    ind = _dev.ind_dev(indsino, (nr,))
    rv = _dev.to_dev(ray_vectors, torch.float32, (nr, 6))
    sino = torch.empty((nr,), dtype=torch.float32, device=rv.device)
    quadric.project_rays(t, sino, ind, rv, nr, 0)
    return _dev.out(sino, host)
