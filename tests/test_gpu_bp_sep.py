"""The column-separable route of the fast adjoint (views whose detector rows
run along z) against the FP64 oracle's exact adjoint (view_proj3D.f:
1096-1370), against the brick kernel's adjoint, and against the fast forward
by the dot-product test.

The route weighs each sample by its z coordinate in exact arithmetic (prefix
sums over a column's rows), where the fast forward rounds that coordinate to
FP32 sample by sample: each z weight differs by at most half an ulp of the
coordinate (below 2^-24 * nz), with signs that vary from sample to sample.
A row within the FP32 error of the cell-boundary search from a boundary
(about 2e-5 of the rows) may be taken in the neighbouring cell, with its
weights extended past [0, 1] by that error (below 1e-3 of a cell); at the
volume's two z faces, where the forward cuts rays off, the cut is the
forward's own, row for row, which the per-voxel comparison with the brick
kernel (the transpose of the forward's own sampling) checks: one row cut
wrongly moves a face voxel by a whole sample, above 1e-2 of its value here.
So the route is held to the FP64 oracle (relative L2 <= 2e-5; it measures
about 2e-6, closer than the brick kernel, whose z weights carry the FP32
rounding) and to the brick adjoint (relative L2 <= 2e-5, the same bar; up to
8e-6 at nz = 256), and its dot-product gap is bounded by
1e-6 of the absolute bilinear form <A|x|, |y|> for Gaussian x, y (it is
about 2e-8 of it) and by 1e-6 relative for positive x, y.

Cases: volume edges that cut the 8 x 8 tiles, 600 views (more than one
launch takes), z sizes that leave lanes
without cells (nz = 70, 130, 256), rows stored in reverse (e_v = -z), a
detector shifted so columns and rows miss the volume, indicator holes and
zero rays, view subsets, and scans that mix separable views with views
whose detector is tilted (those go through the brick kernel and are added).
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


@pytest.fixture(scope="module")
def o64():
    return oracle.restated("f64")


def _case(shape=(37, 29, 70), ns=9, nu=50, nv=44, du=1.6, dv=1.6, u_shift=0.0, v_shift=0.0,
          tilt_every=0, flip_v=False, seed=0, holes=True, R=300.0, sdd=600.0):
    fv = configs3D.frame_vectors("circular_conebeam", {"radius": R, "source_to_detector": sdd},
                                 ns, 2 * np.pi, 0.1, 0).astype(np.float64)
    if flip_v:
        fv[:, 12:15] *= -1.0
    if tilt_every:
        for p in range(0, ns, tilt_every):
            # e_v leans 2 degrees towards e_u: rows no longer run along z
            eu, ev = fv[p, 9:12].copy(), fv[p, 12:15].copy()
            a = np.deg2rad(2.0)
            fv[p, 12:15] = np.cos(a) * ev + np.sin(a) * eu
            fv[p, 9:12] = np.cos(a) * eu - np.sin(a) * ev
    fv = fv.astype(np.float32)
    nx, ny, nz = shape
    geo = (ns, nu, nv, du, dv, -nu * du / 2 + u_shift, -nv * dv / 2 + v_shift)
    g2 = (1.0, 1.0, 1.0, -nx / 2, -ny / 2, -nz / 2, nx, ny, nz)
    rng = np.random.default_rng(seed)
    ind = (rng.random((ns, nu, nv)) > (0.05 if holes else -1)).astype(np.int8)
    return fv, geo, g2, ind, rng


def _run(fv, geo, g2, ind, y, view0=0, nviews=0, vstride=1, x=None, brick=None):
    """A^T y through ct2_backproject (the separable route unless a brick is
    named) and, given x, A x through ct2_project."""
    from ct2 import _dev
    from ct2._lib import CT2_ARITH_FAST, call, ct2_desc
    ns, nu, nv, du, dv, u0, v0 = geo
    dx, dy, dz, x0, y0, z0, nx, ny, nz = g2
    d = ct2_desc()
    call("ct2_desc_init", C.byref(d), ns, nu, nv, du, dv, u0, v0, nx, ny, nz, dx, dy, dz, x0,
         y0, z0)
    d.arith = CT2_ARITH_FAST
    d.view0, d.nviews, d.vstride = view0, nviews, vstride
    if brick is not None:
        d.brick[0], d.brick[1], d.brick[2] = brick
    yd, indd, fvd = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (y, ind, fv))
    out = torch.full((nx, ny, nz), float("nan"), dtype=torch.float32, device="cuda")
    s = _dev.stream_handle()
    call("ct2_backproject", C.byref(d), _dev.ptr(out), _dev.ptr(yd), _dev.ptr(indd),
         _dev.ptr(fvd), s)
    Ax = None
    if x is not None:
        xd = torch.from_numpy(x).cuda()
        Ax = torch.zeros_like(yd)
        call("ct2_project", C.byref(d), _dev.ptr(Ax), _dev.ptr(indd), _dev.ptr(fvd),
             _dev.ptr(xd), s)
        Ax = Ax.cpu().numpy()
    torch.cuda.synchronize()
    return out.cpu().numpy(), Ax


def _oracle(o64, fv, geo, g2, ind, y, views):
    ns = geo[0]
    keep = np.zeros(ns, bool)
    keep[views] = True
    yy = np.where(keep[:, None, None], y, 0).astype(np.float32)
    return o64.view_proj3D.sinobackproject_linear(yy, ind, fv, *geo, *g2)


CASES = {
    "ragged_edges": dict(),
    "nz130_reversed_rows": dict(shape=(24, 33, 130), nv=90, flip_v=True),
    "nz256": dict(shape=(16, 12, 256), nu=30, nv=170, dv=1.6, R=500.0, sdd=1000.0),
    "shifted_detector": dict(u_shift=14.0, v_shift=-20.0),
    "mixed_tilted_views": dict(ns=10, tilt_every=3),
    "nz2": dict(shape=(20, 18, 2), nv=8, dv=0.5),
    # more views than one launch takes (512): launch pairs that store, then add
    "views600": dict(shape=(14, 12, 20), ns=600, nu=18, nv=24, tilt_every=97),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_separable_adjoint_vs_fp64(o64, name):
    fv, geo, g2, ind, rng = _case(**CASES[name])
    y = rng.standard_normal(geo[:3]).astype(np.float32)
    y[:, :, 3] = 0.0   # zero rays
    got, _ = _run(fv, geo, g2, ind, y)
    want = _oracle(o64, fv, geo, g2, ind, y, np.arange(geo[0]))
    assert np.isfinite(got).all()
    assert np.abs(want).max() > 0.1
    e = rel_l2(got, want)
    print(f"{name}: rel L2 vs FP64 {e:.2e}")
    assert e <= 2e-5, e


@pytest.mark.parametrize("view0,nviews,vstride", [(1, 0, 3), (0, 4, 2), (5, 1, 1)])
def test_separable_adjoint_view_subsets(o64, view0, nviews, vstride):
    fv, geo, g2, ind, rng = _case(ns=12, tilt_every=5)
    y = rng.standard_normal(geo[:3]).astype(np.float32)
    got, _ = _run(fv, geo, g2, ind, y, view0, nviews, vstride)
    views = np.arange(view0, geo[0], vstride)
    views = views if nviews == 0 else views[:nviews]
    want = _oracle(o64, fv, geo, g2, ind, y, views)
    assert rel_l2(got, want) <= 2e-5


def _dot(a, b):
    return float(np.dot(a.ravel().astype(np.float64), b.ravel().astype(np.float64)))


@pytest.mark.parametrize("name", ["ragged_edges", "nz130_reversed_rows", "mixed_tilted_views"])
def test_separable_adjoint_dot_product(name):
    fv, geo, g2, ind, rng = _case(**CASES[name])
    x = rng.standard_normal(g2[-3:]).astype(np.float32)
    y = rng.standard_normal(geo[:3]).astype(np.float32)
    Aty, Ax = _run(fv, geo, g2, ind, y, x=x)
    _, Aabs = _run(fv, geo, g2, ind, np.abs(y), x=np.abs(x))
    gap = abs(_dot(Ax, y) - _dot(x, Aty))
    form = _dot(Aabs, np.abs(y))
    print(f"{name}: dot-product gap {gap / form:.2e} of <A|x|,|y|>")
    assert gap <= 1e-6 * form
    # positive vectors: no cancellation in the sums
    xp = rng.uniform(0.5, 1.5, g2[-3:]).astype(np.float32)
    yp = rng.uniform(0.5, 1.5, geo[:3]).astype(np.float32)
    Atyp, Axp = _run(fv, geo, g2, ind, yp, x=xp)
    lhs, rhs = _dot(Axp, yp), _dot(xp, Atyp)
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)


@pytest.mark.parametrize("name", sorted(CASES))
def test_separable_adjoint_vs_brick(name):
    """The separable route and the brick kernel (named brick) agree."""
    fv, geo, g2, ind, rng = _case(**CASES[name])
    y = rng.standard_normal(geo[:3]).astype(np.float32)
    sep, _ = _run(fv, geo, g2, ind, y)
    brk, _ = _run(fv, geo, g2, ind, y, brick=(10, 10, 24))
    e = rel_l2(sep, brk)
    print(f"{name}: rel L2 separable vs brick {e:.2e}")
    assert e <= 2e-5, e


# the cone covers more than the volume's z extent on every plane, so rays
# cross both z faces; the first case has the magnitudes of a 256^3 scan (4.7e7
# face cuts, of which the FP32 crossing alone would misplace hundreds)
FACE_CASES = {
    "faces_256x256x24_360views": dict(shape=(256, 256, 24), ns=360, nu=256, nv=256, du=1.6,
                                      dv=1.6, R=500.0, sdd=1000.0),
    "faces_reversed_rows": dict(shape=(30, 26, 24), ns=40, nu=44, nv=64, dv=0.9, flip_v=True),
    "faces_shifted": dict(shape=(31, 27, 25), ns=40, nu=44, nv=70, dv=0.9, v_shift=-9.0),
}


@pytest.mark.parametrize("name", sorted(FACE_CASES))
def test_separable_adjoint_faces_per_voxel(name):
    """Per voxel, against the brick kernel: the z faces are cut where the
    forward cuts them (positive data, so no voxel is a cancelling sum)."""
    fv, geo, g2, ind, rng = _case(**FACE_CASES[name])
    y = rng.uniform(0.5, 1.5, geo[:3]).astype(np.float32)
    sep, _ = _run(fv, geo, g2, ind, y)
    brk, _ = _run(fv, geo, g2, ind, y, brick=(10, 10, 24))
    scale = np.abs(brk).max()
    err = np.abs(sep - brk)
    faces = max(err[:, :, 0].max(), err[:, :, -1].max()) / scale
    print(f"{name}: max voxel error vs brick {err.max() / scale:.2e}, on the z faces {faces:.2e}")
    assert np.abs(brk[:, :, 0]).max() > 0.1 * scale and np.abs(brk[:, :, -1]).max() > 0.1 * scale
    assert err.max() <= 1e-4 * scale


def test_separable_adjoint_is_deterministic():
    fv, geo, g2, ind, rng = _case(shape=(40, 40, 64), ns=24, nu=64, nv=48, tilt_every=7)
    y = rng.standard_normal(geo[:3]).astype(np.float32)
    a, _ = _run(fv, geo, g2, ind, y)
    b, _ = _run(fv, geo, g2, ind, y)
    assert np.array_equal(a, b)
