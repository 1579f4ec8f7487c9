"""SART composed from the pinned Joseph forward projector and its adjoint --
TEST INFRASTRUCTURE ONLY (the checker of ct2_sart_step and the batched SART;
never imported by the product).

The reference has no SART: its iterative steps are sequential ART
(emil_code/view_proj3D.f:1950-2271) and MLEM (:2507-2831).  SART over the
view subset S (views view0 + j*vstride) is the simultaneous update

    x <- x + lam * A_S^T r / A_S^T m,   r = (b - A_S x) / A_S 1 on used rays,

with A_S the forward projector sinoproj (view_proj3D.f:703-1092) restricted
to S, A_S^T its exact adjoint sinobackproject (:1096-1595), m the mask of
used rays (indicator == 1) in S, and two guards: a ray whose row sum A_S 1
is 0 (it misses the volume) contributes r = 0, and a voxel whose column sum
A_S^T m is 0 (no used ray of S touches it) keeps its value.  Every operator
call goes through the restated C oracle, so the composition inherits the
pins of the forward and the adjoint.
"""
from __future__ import annotations

import numpy as np


def subset_views(ns, view0=0, nviews=0, vstride=1):
    """Views view0 + j*vstride, j < nviews (nviews 0: every such view < ns)."""
    vstride = max(1, vstride)
    views = np.arange(view0, ns, vstride)
    return views if nviews <= 0 else views[:nviews]


def sart_step(o, interp, zonly, lam, sinomat, indsino, fv, ns, nu, nv, du, dv, u0, v0, smat,
              dx, dy, dz, x0, y0, z0, nx, ny, nz, view0=0, nviews=0, vstride=1):
    """One SART update over a view subset; ``o`` is oracle.restated_raw(prec)
    (its precision is the arithmetic of the whole step).  Returns the new
    volume, smat untouched."""
    views = subset_views(ns, view0, nviews, vstride)
    k = views.size
    dt = o.dt
    b = np.asarray(sinomat, dt).reshape(ns, nu, nv)[views]
    ind = (np.ones((ns, nu, nv), np.int8) if indsino is None
           else np.asarray(indsino, np.int8).reshape(ns, nu, nv))[views]
    fvs = np.asarray(fv, np.float32).reshape(ns, 18)[views]
    x = np.asarray(smat, dt).reshape(nx, ny, nz)
    geo = (k, nu, nv, du, dv, u0, v0)
    g2 = (dx, dy, dz, x0, y0, z0, nx, ny, nz)
    ones = np.ones((nx, ny, nz), dt)
    row = o.sinoproj(interp, zonly, ind, fvs, *geo, ones, *g2)
    ax = o.sinoproj(interp, zonly, ind, fvs, *geo, x, *g2)
    used = (ind == 1) & (row > 0)
    r = np.where(used, (b - ax) / np.where(used, row, 1), 0).astype(dt)
    corr = o.sinobackproject(interp, zonly, r, ind, fvs, *geo, *g2)
    col = o.sinobackproject(interp, zonly, (ind == 1).astype(dt), ind, fvs, *geo, *g2)
    return np.where(col > 0, x + lam * corr / np.where(col > 0, col, 1), x).astype(dt)


def sart_sweeps(o, interp, zonly, lam, niter, nsub, sinomat, indsino, fv, ns, nu, nv, du, dv,
                u0, v0, smat, dx, dy, dz, x0, y0, z0, nx, ny, nz):
    """niter passes over nsub interleaved subsets (subset s = views s, s+nsub,
    ...), the order of ct2_batch_sart."""
    x = np.array(smat, o.dt)
    for _ in range(niter):
        for s in range(min(nsub, ns)):
            x = sart_step(o, interp, zonly, lam, sinomat, indsino, fv, ns, nu, nv, du, dv, u0,
                          v0, x, dx, dy, dz, x0, y0, z0, nx, ny, nz, view0=s, vstride=nsub)
    return x
