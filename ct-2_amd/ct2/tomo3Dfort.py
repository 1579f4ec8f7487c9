"""Drop-in for the reference's f2py module ``tomo3Dfort``
(emil_code/tomo3DfortranAux.f, built by emil_code/Makefile as ``tomo3Dfort``):
the voxel embedders (``ct2_phantom_embed``) and the ray-bunching routines
(``ct2_sinoscale``, ``ct2_sinoave``, ``ct2_sinodist``) on the current HIP
device.  numpy in gives numpy out, as the f2py wrappers do; device tensors
stay on the device (the embedders always return numpy)."""
from __future__ import annotations

import torch

from . import _dev, quadric
from ._lib import call


def _embed(t, nx, ny, nz, dx, dy, dz, x0i, y0i, z0i):
    vol = torch.empty((nx, ny, nz), dtype=torch.float32, device=_dev.device())
    quadric.embed(t, vol, nx, ny, nz, dx, dy, dz, x0i, y0i, z0i, 0)
    return vol.cpu().numpy()


def ellipsoid_embed(nx, ny, nz, dx, dy, dz, x0i, y0i, z0i, att, alpha, beta, x0, y0, z0,
                    ax, ay, az):
    """tomo3DfortranAux.f:212-267: att at voxel centres strictly inside the
    ellipsoid, 0 elsewhere; called from emil_code/tomo3D.py:2758."""
    t = quadric.Tables()
    t.add_ellipsoid(att, ax, ay, az, x0, y0, z0, alpha, beta)
    return _embed(t, nx, ny, nz, dx, dy, dz, x0i, y0i, z0i)


def gen_obj_embed(nx, ny, nz, dx, dy, dz, x0i, y0i, z0i, att, nsurf, *surf):
    """tomo3DfortranAux.f:420-582; called from emil_code/tomo3D.py:2106."""
    t = quadric.tables_from_f2py(att, nsurf, *surf)
    return _embed(t, nx, ny, nz, dx, dy, dz, x0i, y0i, z0i)


def ellipse_voxval(x, y, z, att, alpha, beta, x0, y0, z0, ax, ay, az):
    """tomo3DfortranAux.f:179-210: att when (x, y, z) is strictly inside the
    ellipsoid, else 0; called from emil_code/tomo3D.py:2728.  Arrays of points
    (broadcast together) give an array, in one launch."""
    t = quadric.Tables()
    t.add_ellipsoid(att, ax, ay, az, x0, y0, z0, alpha, beta)
    return quadric.voxval_at(t, x, y, z)


def gen_obj_voxval(x, y, z, att, nsurf, *surf):
    """tomo3DfortranAux.f:311-416: att times the product of the surfaces'
    steps at (x, y, z), surface frames by framer; called from
    emil_code/tomo3D.py:2207.  Arrays of points give an array."""
    return quadric.voxval_at(quadric.tables_from_f2py_angles(att, nsurf, *surf), x, y, z)


# ---- ray bunching (tomo3DfortranAux.f:1-176) ----
def sinoscale(sinomat, frame_vectors, ns, nu, nv, du, dv, u0, v0):
    """tomo3DfortranAux.f:1-67 (intent(in,out) sinomat): each bin divided by
    its ray's source-to-bin length; called from emil_code/tomo3D.py:398."""
    from .view_proj3D import _inout
    host = _dev.is_host(sinomat, frame_vectors)
    s = _inout(sinomat, (ns, nu, nv), host)
    fv = _dev.to_dev(frame_vectors, torch.float32, (ns, 18))
    call("ct2_sinoscale", _dev.ptr(s), _dev.ptr(fv), ns, nu, nv, du, dv, u0, v0,
         _dev.stream_handle())
    return _dev.out(s, host)


def sinoave(sinomat, ns, nu, nv, nss, nus, nvs, nds, ndu, ndv, nws, nwu, nwv):
    """tomo3DfortranAux.f:71-122; called from emil_code/tomo3D.py:518."""
    host = _dev.is_host(sinomat)
    s = _dev.to_dev(sinomat, torch.float32, (ns, nu, nv))
    out = torch.empty((nss, nus, nvs), dtype=torch.float32, device=s.device)
    call("ct2_sinoave", _dev.ptr(s), ns, nu, nv, _dev.ptr(out), nss, nus, nvs, nds, ndu, ndv,
         nws, nwu, nwv, _dev.stream_handle())
    return _dev.out(out, host)


def sinodist(sinoin, ns, nu, nv, nss, nus, nvs, nds, ndu, ndv, nws, nwu, nwv):
    """tomo3DfortranAux.f:125-176; called from emil_code/tomo3D.py:534."""
    host = _dev.is_host(sinoin)
    src = _dev.to_dev(sinoin, torch.float32, (nss, nus, nvs))
    out = torch.empty((ns, nu, nv), dtype=torch.float32, device=src.device)
    call("ct2_sinodist", _dev.ptr(out), ns, nu, nv, _dev.ptr(src), nss, nus, nvs, nds, ndu, ndv,
         nws, nwu, nwv, _dev.stream_handle())
    return _dev.out(out, host)
