"""Drop-in for the reference's f2py module ``view_proj3D``
(emil_code/view_proj3D.f, imported at emil_code/tomo3D.py:11).

Same function names, positional signatures and return values as the f2py
wrappers (hidden intent(out) arrays are returned); the work runs in
libct2.so on the current HIP device.  Arrays may be numpy (copied to the
device and back) or torch device tensors (stay on the device).
"""
from __future__ import annotations

import torch

from . import _dev
from ._lib import call

_V = ("ns", "nu", "nv", "du", "dv", "u0", "v0")


def _fwd(sym, indsino, frame_vectors, ns, nu, nv, du, dv, u0, v0, smat,
         dx, dy, dz, x0, y0, z0, nx, ny, nz):
    host = _dev.is_host(indsino, frame_vectors, smat)
    ind = _dev.ind_dev(indsino, (ns, nu, nv))
    fv = _dev.to_dev(frame_vectors, torch.float32, (ns, 18))
    vol = _dev.to_dev(smat, torch.float32, (nx, ny, nz))
    sino = torch.empty((ns, nu, nv), dtype=torch.float32, device=vol.device)
    call(sym, _dev.ptr(sino), _dev.ptr(ind), _dev.ptr(fv), ns, nu, nv, du, dv, u0, v0,
         _dev.ptr(vol), dx, dy, dz, x0, y0, z0, nx, ny, nz, _dev.stream_handle())
    return _dev.out(sino, host)


def _bp(sym, sinomat, indsino, frame_vectors, ns, nu, nv, du, dv, u0, v0,
        dx, dy, dz, x0, y0, z0, nx, ny, nz):
    host = _dev.is_host(sinomat, indsino, frame_vectors)
    ind = _dev.ind_dev(indsino, (ns, nu, nv))
    fv = _dev.to_dev(frame_vectors, torch.float32, (ns, 18))
    sino = _dev.to_dev(sinomat, torch.float32, (ns, nu, nv))
    vol = torch.empty((nx, ny, nz), dtype=torch.float32, device=sino.device)
    call(sym, _dev.ptr(sino), _dev.ptr(ind), _dev.ptr(fv), ns, nu, nv, du, dv, u0, v0,
         _dev.ptr(vol), dx, dy, dz, x0, y0, z0, nx, ny, nz, _dev.stream_handle())
    return _dev.out(vol, host)


def _inout(smat, shape, host):
    """intent(in,out) volume: updated in place when it is already a float32
    contiguous device tensor, otherwise a device copy is updated and returned."""
    if not host and isinstance(smat, torch.Tensor) and smat.is_cuda and \
            smat.dtype == torch.float32 and smat.is_contiguous():
        return smat.view(shape)
    return _dev.to_dev(smat, torch.float32, shape).clone()


def _art(sym, reg_fac, sinomat, indsino, frame_vectors, ns, nu, nv, du, dv, u0, v0, smat,
         dx, dy, dz, x0, y0, z0, nx, ny, nz):
    host = _dev.is_host(sinomat, indsino, frame_vectors, smat)
    ind = _dev.ind_dev(indsino, (ns, nu, nv))
    fv = _dev.to_dev(frame_vectors, torch.float32, (ns, 18))
    sino = _dev.to_dev(sinomat, torch.float32, (ns, nu, nv))
    vol = _inout(smat, (nx, ny, nz), host)
    call(sym, reg_fac, _dev.ptr(sino), _dev.ptr(ind), _dev.ptr(fv), ns, nu, nv, du, dv, u0, v0,
         _dev.ptr(vol), dx, dy, dz, x0, y0, z0, nx, ny, nz, _dev.stream_handle())
    return _dev.out(vol, host)


def _em(sym, sinomat, indsino, frame_vectors, ns, nu, nv, du, dv, u0, v0, smat,
        dx, dy, dz, x0, y0, z0, nx, ny, nz):
    host = _dev.is_host(sinomat, indsino, frame_vectors, smat)
    ind = _dev.ind_dev(indsino, (ns, nu, nv))
    fv = _dev.to_dev(frame_vectors, torch.float32, (ns, 18))
    sino = _dev.to_dev(sinomat, torch.float32, (ns, nu, nv))
    vol = _inout(smat, (nx, ny, nz), host)
    call(sym, _dev.ptr(sino), _dev.ptr(ind), _dev.ptr(fv), ns, nu, nv, du, dv, u0, v0,
         _dev.ptr(vol), dx, dy, dz, x0, y0, z0, nx, ny, nz, _dev.stream_handle())
    return _dev.out(vol, host)


def sinoproj_linear(*a):
    """view_proj3D.f:703-910; called from emil_code/tomo3D.py:1164."""
    return _fwd("ct2_sinoproj_linear", *a)


def sinoproj_nearest(*a):
    """view_proj3D.f:915-1092."""
    return _fwd("ct2_sinoproj_nearest", *a)


def sinobackproject_linear(*a):
    """view_proj3D.f:1096-1370 (exact adjoint); called from emil_code/tomo3D.py:467."""
    return _bp("ct2_sinobackproject_linear", *a)


def sinobackproject_nearest(*a):
    """view_proj3D.f:1374-1595."""
    return _bp("ct2_sinobackproject_nearest", *a)


def pd_backproj_linear(*a):
    """view_proj3D.f:1769-1944 (voxel driven, unweighted)."""
    return _bp("ct2_pd_backproj_linear", *a)


def pd_backproj_nearest(*a):
    """view_proj3D.f:1598-1766."""
    return _bp("ct2_pd_backproj_nearest", *a)


def artstep_linear(*a):
    """view_proj3D.f:1950-2271 (sequential Kaczmarz); called from emil_code/tomo3D.py:1271."""
    return _art("ct2_artstep_linear", *a)


def artstep_nearest(*a):
    """view_proj3D.f:2275-2503."""
    return _art("ct2_artstep_nearest", *a)


def emstep_linear(*a):
    """view_proj3D.f:2507-2831 (MLEM); called from emil_code/tomo3D.py:1430."""
    return _em("ct2_emstep_linear", *a)


def emstep_nearest(*a):
    """view_proj3D.f:2834-3075."""
    return _em("ct2_emstep_nearest", *a)


# ---- analytic phantoms (view_proj3D.f:11-603) ----
def gen_shape_proj(indsino, frame_vectors, nu, nv, du, dv, u0, v0, att, nsurf, *surf):
    """view_proj3D.f:175-455: projection of one general quadric object onto
    one view; called from emil_code/tomo3D.py:2370, :2536.  surf are the
    f2py arguments after nsurf (xu1 .. z0, nplane, nellipsoid, ax_ell, ay_ell,
    az_ell, ncyl, ax_cyl, ay_cyl, ncone, ax_cone, ay_cone)."""
    from . import quadric
    host = _dev.is_host(indsino, frame_vectors)
    t = quadric.tables_from_f2py(att, nsurf, *surf)
    ind = _dev.ind_dev(indsino, (1, nu, nv))
    fv = _dev.to_dev(frame_vectors, torch.float32, (1, 18))
    sino = torch.empty((nu, nv), dtype=torch.float32, device=fv.device)
    quadric.project_views(t, sino, ind, fv, 1, nu, nv, du, dv, u0, v0, 0)
    return _dev.out(sino, host)


def ellipsoid_proj(indsino, frame_vectors, nu, nv, du, dv, u0, v0, ax, ay, az, x0, y0, z0,
                   alpha, beta, att):
    """view_proj3D.f:483-603: chord lengths through a rotated ellipsoid on one
    view; called from emil_code/tomo3D.py:2829, :2866."""
    from . import quadric
    host = _dev.is_host(indsino, frame_vectors)
    t = quadric.Tables()
    t.add_ellipsoid(att, ax, ay, az, x0, y0, z0, alpha, beta)
    ind = _dev.ind_dev(indsino, (1, nu, nv))
    fv = _dev.to_dev(frame_vectors, torch.float32, (1, 18))
    sino = torch.empty((nu, nv), dtype=torch.float32, device=fv.device)
    quadric.project_views(t, sino, ind, fv, 1, nu, nv, du, dv, u0, v0, 0)
    return _dev.out(sino, host)
