"""Drop-in for the reference's ``FDK`` module (emil_code/FDK.py) used by
emil_code/recon3D.py:65-131.  Same function names and arguments; the work runs
in libct2.so.  Arrays may be numpy (host round trip) or torch device tensors.
The reference computes in float64 numpy; these run in FP32 on the device
with FP64 weights and filter coefficients."""
from __future__ import annotations

import numpy as np
import torch

from . import _dev
from ._lib import call


def _grid_params(v):
    """(start, step, offset) of an evenly spaced coordinate vector."""
    v = np.asarray(v, np.float64).ravel()
    if v.size == 1:
        return float(v[0]), 1.0
    return float(v[0]), float(v[1] - v[0])


def fdk_sino_weight3D(radius, center2detector, ns, vu, vv, sinomat):
    """FDK.py:10-23.  vu, vv are the (nu, nv) detector coordinate meshgrids
    (recon3D.py:59-62); they must be evenly spaced grids, as there."""
    host = _dev.is_host(sinomat)
    vu = np.asarray(vu, np.float64)
    vv = np.asarray(vv, np.float64)
    nu, nv = vu.shape
    s = _dev.to_dev(sinomat, torch.float32, (ns, nu, nv))
    u00, du = _grid_params(vu[:, 0])
    v00, dv = _grid_params(vv[0, :])
    out = torch.empty_like(s)
    call("ct2_fdk_sino_weight3d", radius, center2detector, ns, nu, nv, u00, du, 0.0, v00, dv, 0.0,
         _dev.ptr(s), _dev.ptr(out), _dev.stream_handle())
    return _dev.out(out, host)


def fdk_sino_filter3D(dp, dim, sinomat, opt_window='', L=1):
    """FDK.py:25-103: ramp filter along axis dim (1 = u, 2 = v)."""
    host = _dev.is_host(sinomat)
    s = _dev.to_dev(sinomat, torch.float32)
    ns, nu, nv = s.shape
    out = torch.empty_like(s)
    call("ct2_fdk_sino_filter3d", dp, dim, ns, nu, nv, _dev.ptr(s), _dev.ptr(out),
         1 if opt_window == 'Hann' else 0, float(L), _dev.stream_handle())
    return _dev.out(out, host)


def fdk_backproj_linear3D(ns, nu, nv, du, dv, u0, v0, radius, source2detector, uoffset, voffset,
                          vs, vx, vy, vz, sinomat):
    """FDK.py:130-208.  vs (angles), vx, vy (flattened meshgrid of x, y as in
    recon3D.py:111-115) and vz must be evenly spaced.  Returns the reference's
    recon[ny*nx, nz] layout (row = iy*nx + ix), scaled by pi/ns."""
    cube = fdk_backproj_cube(ns, nu, nv, du, dv, u0, v0, radius, source2detector, uoffset,
                             voffset, vs, vx, vy, vz, sinomat)
    host = _dev.is_host(sinomat)
    c = torch.as_tensor(cube) if host else cube
    nx, ny, nz = c.shape
    r = c.permute(1, 0, 2).reshape(ny * nx, nz)
    return r.numpy() if host else r


def fdk_backproj_cube(ns, nu, nv, du, dv, u0, v0, radius, source2detector, uoffset, voffset,
                      vs, vx, vy, vz, sinomat):
    """fdk_backproj_linear3D returning the volume cube[ix, iy, iz]
    (the reshaping of recon3D.py:134-139 already applied)."""
    host = _dev.is_host(sinomat)
    vs = np.asarray(vs, np.float64).ravel()
    vx = np.asarray(vx, np.float64).ravel()
    vy = np.asarray(vy, np.float64).ravel()
    vz = np.asarray(vz, np.float64).ravel()
    xs = np.unique(vx)
    ys = np.unique(vy)
    nx, ny, nz = xs.size, ys.size, vz.size
    x00, dx = _grid_params(xs)
    y00, dy = _grid_params(ys)
    z00, dz = _grid_params(vz)
    s0, ds = _grid_params(vs)
    s = _dev.to_dev(sinomat, torch.float32, (ns, nu, nv))
    rec = torch.empty((nx, ny, nz), dtype=torch.float32, device=s.device)
    call("ct2_fdk_backproj_linear3d", ns, nu, nv, du, dv, u0, v0, radius, source2detector,
         uoffset, voffset, s0, ds, nx, ny, nz, x00, y00, z00, dx, dy, dz, 0.0, 0.0, 0.0,
         _dev.ptr(s), _dev.ptr(rec), _dev.stream_handle())
    return _dev.out(rec, host)


def fdk_reconstruct(sinomat, sino, image, opt_window='', L=1):
    """The FDK pipeline of emil_code/recon3D.py:50-139 for a circular
    cone-beam tomo3D.sinogram3D ``sino`` onto the voxel centres of
    tomo3D.image3D ``image``: angles vs = s0 + p*ds, detector and voxel
    offsets 0.5, weighting called with (radius, source_to_detector) as at
    recon3D.py:65-67, ramp filter along u (dim 1, pitch du), then
    linear-interpolation backprojection.  Returns the (nx, ny, nz) cube."""
    ns, nu, nv = sino.ns, sino.nu, sino.nv
    R = sino.parms["radius"]
    SDD = sino.parms["source_to_detector"]
    vs = sino.s0 + np.arange(ns) * sino.ds
    vu = sino.u0 + (np.arange(nu) + 0.5) * sino.du
    vv = sino.v0 + (np.arange(nv) + 0.5) * sino.dv
    vvg, vug = np.meshgrid(vv, vu)
    w = fdk_sino_weight3D(R, SDD, ns, vug, vvg, sinomat)
    f = fdk_sino_filter3D(sino.du, 1, w, opt_window, L)
    vx = image.x0 + (np.arange(image.nx) + 0.5) * image.dx
    vy = image.y0 + (np.arange(image.ny) + 0.5) * image.dy
    vz = image.z0 + (np.arange(image.nz) + 0.5) * image.dz
    vxg, vyg = np.meshgrid(vx, vy)
    return fdk_backproj_cube(ns, nu, nv, sino.du, sino.dv, sino.u0, sino.v0, R, SDD, 0.5, 0.5,
                             vs, vxg.ravel(), vyg.ravel(), vz, f)
