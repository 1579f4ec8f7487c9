"""ct2_backproject_ws: the adjoint with a caller-owned workspace for the
column-separable route (a header and one prefix-sum table per view and
detector column, built once per call by k_sep_tables).

* A workspace of ct2_backproject_work_bytes gives the same bits as
  ct2_backproject, which allocates the same workspace on the stream.
* A smaller one processes the views in groups (the x- and y-driven launches
  of each group add in stream order); that only regroups the FP32 sums of
  each voxel, so the result agrees with the one-group adjoint to FP32
  rounding (relative L2 below 1e-6) and with the FP64 oracle as closely.
* Scans that mix separable views with tilted ones (which the brick kernel
  adds after the route) give the same sums through either entry.
* A workspace too small for one view's tables, or not 16-byte aligned, is
  rejected with CT2_EINVAL before any launch; work_bytes is 0 where the
  route does not apply.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

import oracle
from cases import rel_l2
from ct2 import configs3D

pytestmark = pytest.mark.gpu


def _case(ns=24, nu=40, nv=52, shape=(30, 26, 40), tilt_every=0, seed=3):
    fv = configs3D.frame_vectors("circular_conebeam", {"radius": 300.0, "source_to_detector": 600.0},
                                 ns, 2 * np.pi, 0.2, 0).astype(np.float64)
    if tilt_every:
        for p in range(0, ns, tilt_every):
            eu, ev = fv[p, 9:12].copy(), fv[p, 12:15].copy()
            a = np.deg2rad(2.0)
            fv[p, 12:15] = np.cos(a) * ev + np.sin(a) * eu
            fv[p, 9:12] = np.cos(a) * eu - np.sin(a) * ev
    fv = fv.astype(np.float32)
    nx, ny, nz = shape
    geo = (ns, nu, nv, 1.6, 1.6, -nu * 0.8, -nv * 0.8)
    g2 = (1.0, 1.0, 1.0, -nx / 2, -ny / 2, -nz / 2, nx, ny, nz)
    rng = np.random.default_rng(seed)
    y = rng.random((ns, nu, nv)).astype(np.float32) + 0.1
    ind = (rng.random((ns, nu, nv)) > 0.03).astype(np.int8)
    return fv, geo, g2, y, ind


def _desc(geo, g2):
    from ct2._lib import CT2_ARITH_FAST, call, ct2_desc
    ns, nu, nv, du, dv, u0, v0 = geo
    dx, dy, dz, x0, y0, z0, nx, ny, nz = g2
    d = ct2_desc()
    call("ct2_desc_init", C.byref(d), ns, nu, nv, du, dv, u0, v0, nx, ny, nz, dx, dy, dz, x0,
         y0, z0)
    d.arith = CT2_ARITH_FAST
    return d


def _bp(d, g2, y, ind, fv, work=None, nbytes=0, offset=0):
    from ct2 import _dev
    from ct2._lib import call
    nx, ny, nz = g2[6:]
    yd, indd, fvd = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (y, ind, fv))
    out = torch.full((nx, ny, nz), float("nan"), dtype=torch.float32, device="cuda")
    s = _dev.stream_handle()
    if work is None:
        call("ct2_backproject", C.byref(d), _dev.ptr(out), _dev.ptr(yd), _dev.ptr(indd),
             _dev.ptr(fvd), s)
    else:
        call("ct2_backproject_ws", C.byref(d), _dev.ptr(out), _dev.ptr(yd), _dev.ptr(indd),
             _dev.ptr(fvd), C.c_void_p(work.data_ptr() + offset), C.c_size_t(nbytes), s)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _work_bytes(d):
    from ct2._lib import lib
    return int(lib().ct2_backproject_work_bytes(C.byref(d)))


@pytest.mark.parametrize("tilt_every", [0, 5])
def test_workspace_entry_matches_allocating_entry(tilt_every):
    fv, geo, g2, y, ind = _case(tilt_every=tilt_every)
    d = _desc(geo, g2)
    n = _work_bytes(d)
    ns, nu, nv = geo[:3]
    assert n == 256 + ns * nu * (nv + 1) * 16
    work = torch.empty(n, dtype=torch.uint8, device="cuda")
    a = _bp(d, g2, y, ind, fv)
    b = _bp(d, g2, y, ind, fv, work, n)
    assert np.isfinite(a).all()
    assert np.array_equal(a, b)
    # the same buffer again (its header and tables are rewritten by each call)
    c = _bp(d, g2, y, ind, fv, work, n)
    assert np.array_equal(a, c)


@pytest.mark.parametrize("views_per_group", [1, 7])
def test_small_workspace_groups_views(views_per_group):
    fv, geo, g2, y, ind = _case(tilt_every=6)
    d = _desc(geo, g2)
    ns, nu, nv = geo[:3]
    nbytes = 256 + views_per_group * nu * (nv + 1) * 16
    work = torch.empty(nbytes + 64, dtype=torch.uint8, device="cuda")
    a = _bp(d, g2, y, ind, fv)
    b = _bp(d, g2, y, ind, fv, work, nbytes)
    assert np.isfinite(b).all()
    assert rel_l2(b, a) < 1e-6
    o64 = oracle.restated("f64")
    ref = o64.view_proj3D.sinobackproject_linear(y, ind, fv, *geo, *g2)
    assert abs(rel_l2(b, ref) - rel_l2(a, ref)) < 1e-6
    assert rel_l2(b, ref) < 2e-5


def test_workspace_rejections():
    from ct2._lib import CT2Error
    fv, geo, g2, y, ind = _case()
    d = _desc(geo, g2)
    ns, nu, nv = geo[:3]
    work = torch.empty(_work_bytes(d) + 64, dtype=torch.uint8, device="cuda")
    with pytest.raises(CT2Error, match="holds no view"):
        _bp(d, g2, y, ind, fv, work, 256 + nu * (nv + 1) * 16 - 1)
    with pytest.raises(CT2Error, match="16-byte aligned"):
        _bp(d, g2, y, ind, fv, work, _work_bytes(d), offset=4)
    # exact arithmetic or a named brick: no route, no workspace
    from ct2._lib import CT2_ARITH_EXACT
    e = _desc(geo, g2)
    e.arith = CT2_ARITH_EXACT
    assert _work_bytes(e) == 0
    f = _desc(geo, g2)
    f.brick[0], f.brick[1], f.brick[2] = 8, 8, 8
    assert _work_bytes(f) == 0
