"""Weighted ray-bunch and derivative measurements and their ART
(``ct2_stencil_project``, ``ct2_stencil_weights``, ``ct2_stencil_art``).

The reference keeps the call sites and checks of these GPU kernels
(tests/cweighted_art_test.py:80-140: ``cweighted_project_to``,
``put_weights``, ``c_ss_weighted_art_step``; tests/cderiv_art_test.py:106-140:
``get_deriv_weights``, ``cderiv_art_step``) but not their sources.  The
semantics here are the ones those checks state:

* coarse bin (I,J,K) measures ``sum_q w_q (A x)[ray q]`` over the kernel
  ``w[ws][wu][wv]`` placed at fine ray (I*nss, J*nsu, K*nsv) minus the kernel
  centre, i.e. ``scipy.ndimage.convolve(proj, w[::-1,::-1,::-1],
  mode='constant')[::nss, ::nsu, ::nsv]`` (cweighted_art_test.py:92);
* ``put_weights`` gives each used bin's squared row norm |r_b|^2 and puts the
  sum of the used rows in the image (the test's ``sum(image**2)`` against
  ``data.weights``);
* the ART step is Kaczmarz over the used bins in (I,J,K) order,
  ``x += reg_fac (m_b - r_b.x) r_b / |r_b|^2``;
* the derivative data are the kernel ``[-1, +1]`` along u with unit strides
  (``data.mat[:, j] = p[:, j+1] - p[:, j]``, cderiv_art_test.py:109).

A bin is used when the fine indicator at its anchor (I*nss, J*nsu, K*nsv) is
1, as the tests set it.  The CUDA launch-shape arguments of the reference's
methods are accepted and ignored.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _dev
from ._lib import (CT2_ARITH_EXACT, CT2_AXIS_FREE, CT2_AXIS_Z, CT2_F32, CT2_LINEAR, CT2_NEAREST,
                   call, ct2_desc, ct2_stencil)

DERIV_U = np.array([[[-1.0], [1.0]]], np.float32)     # d/du kernel, unit strides


def coarse_shape(ns, nu, nv, nss, nsu, nsv):
    return (-(-ns // nss), -(-nu // nsu), -(-nv // nsv))


def _desc(ns, nu, nv, du, dv, u0, v0, grid, interp, axis):
    dx, dy, dz, x0, y0, z0, nx, ny, nz = grid
    d = ct2_desc()
    call("ct2_desc_init", C.byref(d), ns, nu, nv, du, dv, u0, v0, nx, ny, nz,
         dx, dy, dz, x0, y0, z0)
    d.interp = {"linear": CT2_LINEAR, "nearest": CT2_NEAREST}[interp]
    d.axis = {"free": CT2_AXIS_FREE, "z": CT2_AXIS_Z}[axis]
    d.arith = CT2_ARITH_EXACT
    d.vol_dtype = CT2_F32
    return d


def _stencil(w, nss, nsu, nsv):
    w = np.asarray(w.detach().cpu() if isinstance(w, torch.Tensor) else w, np.float32)
    if w.ndim != 3:
        raise ValueError("stencil weights must be a 3-D array [ws, wu, wv]")
    st = ct2_stencil(*w.shape, int(nss), int(nsu), int(nsv))
    return st, _dev.to_dev(w, torch.float32, w.shape)


def stencil_project(smat, w, nss, nsu, nsv, frame_vectors, ns, nu, nv, du, dv, u0, v0, grid,
                    interp="linear", axis="free"):
    """Coarse measurements [NS, NU, NV] of the volume (cweighted_project_to)."""
    host = _dev.is_host(smat, frame_vectors)
    nx, ny, nz = grid[6:9]
    st, wd = _stencil(w, nss, nsu, nsv)
    vol = _dev.to_dev(smat, torch.float32, (nx, ny, nz))
    fv = _dev.to_dev(frame_vectors, torch.float32, (ns, 18))
    fine = torch.empty((ns, nu, nv), dtype=torch.float32, device=vol.device)
    small = torch.empty(coarse_shape(ns, nu, nv, nss, nsu, nsv), dtype=torch.float32,
                        device=vol.device)
    d = _desc(ns, nu, nv, du, dv, u0, v0, grid, interp, axis)
    call("ct2_stencil_project", C.byref(d), C.byref(st), _dev.ptr(wd), _dev.ptr(vol),
         _dev.ptr(fv), _dev.ptr(fine), _dev.ptr(small), _dev.stream_handle())
    return _dev.out(small, host)


def stencil_weights(w, nss, nsu, nsv, indsino, frame_vectors, ns, nu, nv, du, dv, u0, v0, grid,
                    interp="linear", axis="free"):
    """(weights [NS, NU, NV], image): squared row norms of the used bins and
    the sum of their rows (put_weights / get_deriv_weights)."""
    host = _dev.is_host(indsino, frame_vectors)
    nx, ny, nz = grid[6:9]
    st, wd = _stencil(w, nss, nsu, nsv)
    ind = _dev.ind_dev(indsino, (ns, nu, nv))
    fv = _dev.to_dev(frame_vectors, torch.float32, (ns, 18))
    dev = fv.device
    weights = torch.empty(coarse_shape(ns, nu, nv, nss, nsu, nsv), dtype=torch.float32, device=dev)
    image = torch.zeros((nx, ny, nz), dtype=torch.float32, device=dev)
    scratch = torch.zeros((nx, ny, nz), dtype=torch.float32, device=dev)
    d = _desc(ns, nu, nv, du, dv, u0, v0, grid, interp, axis)
    call("ct2_stencil_weights", C.byref(d), C.byref(st), _dev.ptr(wd), _dev.ptr(ind),
         _dev.ptr(fv), _dev.ptr(weights), _dev.ptr(image), _dev.ptr(scratch),
         _dev.stream_handle())
    return _dev.out(weights, host), _dev.out(image, host)


def stencil_art_step(smat, w, nss, nsu, nsv, small, weights, indsino, frame_vectors, ns, nu, nv,
                     du, dv, u0, v0, grid, reg_fac=1.0, interp="linear", axis="free"):
    """One Kaczmarz sweep over the used coarse bins; returns the volume
    (updated in place when it is a contiguous float32 device tensor)."""
    from .view_proj3D import _inout
    host = _dev.is_host(smat, small, weights, indsino, frame_vectors)
    nx, ny, nz = grid[6:9]
    st, wd = _stencil(w, nss, nsu, nsv)
    cs = coarse_shape(ns, nu, nv, nss, nsu, nsv)
    vol = _inout(smat, (nx, ny, nz), host)
    m = _dev.to_dev(small, torch.float32, cs)
    wt = _dev.to_dev(weights, torch.float32, cs)
    ind = _dev.ind_dev(indsino, (ns, nu, nv))
    fv = _dev.to_dev(frame_vectors, torch.float32, (ns, 18))
    d = _desc(ns, nu, nv, du, dv, u0, v0, grid, interp, axis)
    call("ct2_stencil_art", C.byref(d), C.byref(st), _dev.ptr(wd), C.c_float(reg_fac),
         _dev.ptr(m), _dev.ptr(wt), _dev.ptr(ind), _dev.ptr(fv), _dev.ptr(vol),
         _dev.stream_handle())
    return _dev.out(vol, host)


# ------------------------------------------- the reference's method surface --
def _geo(image, data):
    return (data.frame_vectors, data.ns, data.nu, data.nv, data.du, data.dv, data.u0, data.v0,
            image._grid())


def cweighted_project_to(image, data, rayweights, nss, nsu, nsv, *launch):
    """data.smallsino = weighted ray-bunch projection of image."""
    data.smallsino = stencil_project(image.mat, rayweights, nss, nsu, nsv, *_geo(image, data))


def put_weights(image, data, rayweights, nss, nsu, nsv, *launch):
    """data.weights = squared row norms; image.mat = sum of the used rows."""
    data.weights, image.mat = stencil_weights(rayweights, nss, nsu, nsv, data.indicator,
                                              *_geo(image, data))


def c_ss_weighted_art_step(image, data, rayweights, nss, nsu, nsv, *launch, reg_fac=1.0):
    """ART over the coarse bins against data.smallsino and data.weights."""
    image.mat = stencil_art_step(image.mat, rayweights, nss, nsu, nsv, data.smallsino,
                                 data.weights, data.indicator, *_geo(image, data),
                                 reg_fac=reg_fac)


def get_deriv_weights(image, data, *launch):
    put_weights(image, data, DERIV_U, 1, 1, 1)


def cderiv_art_step(image, data, *launch, reg_fac=1.0):
    """ART against u-differences held in data.mat (cderiv_art_test.py:109)."""
    image.mat = stencil_art_step(image.mat, DERIV_U, 1, 1, 1, data.mat, data.weights,
                                 data.indicator, *_geo(image, data), reg_fac=reg_fac)
