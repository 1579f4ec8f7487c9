"""Iteration drivers composed from the pinned oracle pieces -- TEST
INFRASTRUCTURE ONLY (the checker of ct2.solvers; never imported by the
product).

* ``asd_pocs`` restates the ASD-POCS loop of
  emil_code/examples/jawPhantomCCB.py:191-327 on plain arrays: ``nart`` ART
  sweeps (artstep_linear, view_proj3D.f:1950-2271), zero_edges
  (tomo3D.py:928-943), make_positive (:902-903), the image change adist
  (dist_to, :946-951), gf scaled by adist on the first iteration (:214-215),
  the TV value (total_var3d, totalvar.f:35-62), the data distance over the
  used rays (:319-324) of the forward projection (sinoproj_linear,
  view_proj3D.f:703-910), meps at iteration iter_eps (:250-251), ``ngradmax``
  steps x <- max(x - gf*g, 0) with zero edges, g the normalised TV gradient
  (total_var_grad3d, totalvar.f:476-602; image3D.tv_grad and scale,
  tomo3D.py:1031-1043), then the gf and reg_fac rules (:321-324).  It
  returns the final image and the records (iter, ddist, posttv, ierr) that
  the reference writes to derrEvol.txt.
* ``asd_pocs_sart`` restates ct2.solvers.AsdPocs: the same loop with the
  data step a sweep of ordered-subset SART (oracle/sart.py, nsub interleaved
  view subsets, relaxation reg_fac), the image and the step sizes gf and
  reg_fac in the oracle's precision, and every distance with its difference
  in float64 (ct2_distance).
* ``em_tv`` restates ct2.solvers.em_tv: the loop of examples/arttest.py:110-128
  with its EM line (:117, emstep_linear, view_proj3D.f:2507-2831;
  tomo3D.py:1394-1435) in place of ART, followed by jawPhantomCCB.py's TV
  descent and step rule: per iteration one EM update,
  positivity, adist, gf = alpha*adist on the first iteration, ``n_tv``
  normalised TV-gradient steps with positivity, and gf *= red_fact when the
  descent moved the image at least ``conv`` times as far as the EM step.

Every operator is a method of ``o`` (oracle.restated_raw(prec), whose
forward, ART, EM and TV are pinned against the compiled reference), so the
compositions inherit those pins.
"""
from __future__ import annotations

import numpy as np


def dist_to(a, b, w=None):
    """tomo3D.py:319-324 and :946-951 as numpy evaluates them on float32
    arrays: difference and square in float32, the weights (0/1) and the sum
    in float64."""
    d = np.asarray(a, np.float32).ravel() - np.asarray(b, np.float32).ravel()
    d = (d * d).astype(np.float64)
    if w is not None:
        d = d * np.asarray(w, np.float64).ravel()
    return float(np.sqrt(np.sum(d, dtype=np.float64)))


def dist64(a, b):
    """sqrt(sum (a-b)^2) with the difference in float64 (ct2_distance)."""
    d = np.asarray(a, np.float64).ravel() - np.asarray(b, np.float64).ravel()
    return float(np.sqrt(np.sum(d * d)))


def zero_edges(x):
    """tomo3D.py:928-943: the two-voxel border of every face set to 0."""
    x[0:2] = 0.0
    x[-2:] = 0.0
    x[:, 0:2] = 0.0
    x[:, -2:] = 0.0
    x[:, :, 0:2] = 0.0
    x[:, :, -2:] = 0.0


def make_positive(x):
    """tomo3D.py:902-903."""
    x[x < 0.0] = 0.0


def asd_pocs(o, x, b, ind, fv, geo, g2, max_iter, nart=1, ngradmax=20, gf=0.2, reg_fac=1.0,
             red_fact=0.98, red_reg=0.995, conv=0.99, iter_eps=500, truth=None, lam1=0.0):
    """ASD-POCS (jawPhantomCCB.py:191-327) from image x (not modified) with
    data b; returns (image, records, (gf, reg_fac, meps) after the last
    iteration)."""
    ns, nu, nv, du, dv, u0, v0 = geo
    dx, dy, dz, x0, y0, z0, nx, ny, nz = g2
    dt = o.dt
    x = np.array(x, dt).reshape(nx, ny, nz)
    b = np.asarray(b, dt).reshape(ns, nu, nv)
    used = np.ones((ns, nu, nv), np.int8) if ind is None else np.asarray(ind, np.int8)
    meps = -1.0
    records = []
    for it in range(1, max_iter + 1):
        pre = x.copy()
        for _ in range(nart):
            x = o.artstep(0, 0, reg_fac, b, ind, fv, *geo, x, *g2)
        zero_edges(x)
        make_positive(x)
        adist = dist_to(x, pre)
        if it == 1:
            gf *= adist
        posttv = o.total_var3d(lam1, x, nx, ny, nz)
        work = o.sinoproj(0, 0, ind, fv, *geo, x, *g2)
        ddist = dist_to(b, work, used)
        if it == iter_eps:
            meps = ddist
        ierr = dist_to(x, truth) if truth is not None else float("nan")
        records.append((it, ddist, posttv, ierr))
        pre = x.copy()
        for _ in range(ngradmax):
            g, _ = o.total_var_grad3d(lam1, x, nx, ny, nz, 1)
            # image3D.scale: factor * mat, a float32 array times a scalar
            x = x - (gf * g).astype(dt)
            zero_edges(x)
            make_positive(x)
        gdist = dist_to(x, pre)
        if gdist >= conv * adist and ddist > meps:
            gf *= red_fact
        reg_fac *= red_reg
    return x, records, (gf, reg_fac, meps)


def em_tv(o, x, b, ind, fv, geo, g2, n_em, n_tv=5, eps=1e-8, alpha=0.2, red_fact=0.98,
          conv=0.99):
    """ct2.solvers.em_tv from image x (not modified) with data b, in the
    precision of ``o`` (the image and gf in float32 for the FP32 oracle, as
    the device loop keeps them; float64 throughout for the FP64 oracle);
    returns (image, [(adist, gdist, gf) per iteration])."""
    dx, dy, dz, x0, y0, z0, nx, ny, nz = g2
    ft = o.dt
    x = np.array(x, ft).reshape(nx, ny, nz)
    b = np.asarray(b, ft)
    gf = None
    hist = []
    for _ in range(n_em):
        prev = x.copy()
        x = o.emstep(0, 0, b, ind, fv, *geo, x, *g2).astype(ft)
        make_positive(x)
        adist = dist64(x, prev)
        if gf is None:
            gf = ft(ft(alpha) * ft(adist))
        prev = x.copy()
        for _ in range(n_tv):
            g, _ = o.total_var_grad3d(0.0, x, nx, ny, nz, 1, eps)
            x = x - gf * g.astype(ft)
            make_positive(x)
        gdist = dist64(x, prev)
        hist.append((adist, gdist, float(gf)))
        if gdist >= conv * adist:
            gf = ft(gf * ft(red_fact))
    return x, hist


def asd_pocs_sart(o, x, b, ind, fv, geo, g2, max_iter, nsub=10, ngrad=20, gf=0.2, reg_fac=1.0,
                  red_fact=0.98, red_reg=0.995, conv=0.99, iter_eps=500, truth=None, eps=1e-9):
    """ct2.solvers.AsdPocs from image x (not modified) with data b, in the
    precision of ``o``; returns (image, records, (gf, reg_fac, meps))."""
    from . import sart as osart
    ns, nu, nv, du, dv, u0, v0 = geo
    dx, dy, dz, x0, y0, z0, nx, ny, nz = g2
    ft = o.dt
    x = np.array(x, ft).reshape(nx, ny, nz)
    b = np.asarray(b, ft).reshape(ns, nu, nv)
    used = np.ones((ns, nu, nv), np.float64) if ind is None else np.asarray(ind, np.float64)
    gf, reg_fac = ft(gf), ft(reg_fac)
    meps = -1.0
    records = []
    for it in range(1, max_iter + 1):
        pre = x.copy()
        x = osart.sart_sweeps(o, 0, 0, reg_fac, 1, nsub, b, ind, fv, *geo, x, *g2).astype(ft)
        zero_edges(x)
        make_positive(x)
        adist = dist64(x, pre)
        if it == 1:
            gf = ft(gf * ft(adist))
        tv = o.total_var3d(0.0, x, nx, ny, nz, eps)
        proj = o.sinoproj(0, 0, ind, fv, *geo, x, *g2)
        r = np.asarray(proj, np.float64) - np.asarray(b, np.float64)
        ddist = float(np.sqrt(np.sum(used * r * r)))
        if it == iter_eps:
            meps = ddist
        ierr = dist64(x, truth) if truth is not None else float("nan")
        records.append((it, ddist, tv, ierr))
        pre = x.copy()
        for _ in range(ngrad):
            g, _ = o.total_var_grad3d(0.0, x, nx, ny, nz, 1, eps)
            x = x - gf * g.astype(ft)
            zero_edges(x)
            make_positive(x)
        gdist = dist64(x, pre)
        if gdist >= conv * adist and ddist > meps:
            gf = ft(gf * ft(red_fact))
        reg_fac = ft(reg_fac * ft(red_reg))
    return x, records, (float(gf), float(reg_fac), meps)
