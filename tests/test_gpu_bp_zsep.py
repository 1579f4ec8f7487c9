"""The plane-separable route of the fast adjoint under the z-forced projector
(tomosynthesis views: flat detector parallel to the z planes, e_u along x,
e_v along y; configs3D.py:615-654 with view_proj3D_tomo.f:4-148) against the
brick kernel's adjoint, the FP64 oracle's z-forced adjoint and the fast
forward by the dot-product test.

The route uses the forward's own FP32 cell coordinates, weights and ray
weights (the brick kernel's matrix entries) and differs from the brick
kernel only in how the sums are grouped (FP32 per y resampling, then per
x row, against FP64 LDS sums): within 1e-6 relative L2 of it.  Against the
FP64 oracle the bar is the fast arithmetic's, 2e-5, and the dot-product gap
is bounded by 1e-6 of <A|x|, |y|>.

Cases: volume edges that cut the 8 x 64 x 8 tiles and leave z sizes that are
not multiples of 4, a detector shifted so columns and rows miss the volume,
e_u along -x, a coarse detector (two voxels per bin), indicator holes and
zero rays, view subsets, scans that mix tomosynthesis views with tilted ones
(the brick kernel adds those), a detector finer than the route accepts (the
brick kernel takes every view), and an output that is not 16-byte aligned.
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
    return oracle.restated_raw("f64")


def _case(shape=(45, 70, 13), vox=(0.5, 0.5, 2.0), ns=7, nu=60, nv=150, du=0.5, dv=0.5,
          u_shift=0.0, v_shift=0.0, flip_u=False, tilt_every=0, seed=0, holes=True, R=200.0,
          c2d=10.0):
    fv = configs3D.frame_vectors("tomosynthesis", {"radius": R, "center_to_detector": c2d},
                                 ns, np.deg2rad(50.0), np.deg2rad(65.0), 1).astype(np.float64)
    if flip_u:
        fv[:, 9:12] *= -1.0
    if tilt_every:
        for p in range(0, ns, tilt_every):
            # e_v leans 3 degrees out of the detector plane: not plane-separable
            a = np.deg2rad(3.0)
            fv[p, 12:15] = (0.0, np.cos(a), np.sin(a))
    fv = fv.astype(np.float32)
    nx, ny, nz = shape
    geo = (ns, nu, nv, du, dv, -nu * du / 2 + u_shift, -nv * dv / 2 + v_shift)
    g2 = (vox[0], vox[1], vox[2], -nx * vox[0] / 2, -ny * vox[1] / 2, 0.0, nx, ny, nz)
    rng = np.random.default_rng(seed)
    ind = (rng.random((ns, nu, nv)) > (0.05 if holes else -1)).astype(np.int8)
    return fv, geo, g2, ind, rng


def _desc(geo, g2, brick=None, view0=0, nviews=0, vstride=1):
    from ct2._lib import CT2_ARITH_FAST, CT2_AXIS_Z, call, ct2_desc
    ns, nu, nv, du, dv, u0, v0 = geo
    dx, dy, dz, x0, y0, z0, nx, ny, nz = g2
    d = ct2_desc()
    call("ct2_desc_init", C.byref(d), ns, nu, nv, du, dv, u0, v0, nx, ny, nz, dx, dy, dz, x0,
         y0, z0)
    d.arith = CT2_ARITH_FAST
    d.axis = CT2_AXIS_Z
    d.view0, d.nviews, d.vstride = view0, nviews, vstride
    if brick is not None:
        d.brick[0], d.brick[1], d.brick[2] = brick
    return d


def _adjoint(fv, geo, g2, ind, y, brick=None, offset=0, **sub):
    """A^T y through ct2_backproject with the axis forced to z: the plane-
    separable route unless a brick is named; offset shifts the output by
    that many floats inside a larger buffer."""
    from ct2 import _dev
    from ct2._lib import call
    nx, ny, nz = g2[6:]
    d = _desc(geo, g2, brick, **sub)
    yd, fvd = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (y, fv))
    indd = None if ind is None else torch.from_numpy(ind).cuda()
    buf = torch.full((nx * ny * nz + offset,), float("nan"), dtype=torch.float32, device="cuda")
    out = buf[offset:]
    call("ct2_backproject", C.byref(d), _dev.ptr(out), _dev.ptr(yd),
         None if indd is None else _dev.ptr(indd), _dev.ptr(fvd), _dev.stream_handle())
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(nx, ny, nz)


def _forward(fv, geo, g2, ind, x):
    from ct2 import _dev
    from ct2._lib import call
    d = _desc(geo, g2)
    xd, fvd, indd = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (x, fv, ind))
    out = torch.zeros(geo[:3], dtype=torch.float32, device="cuda")
    call("ct2_project", C.byref(d), _dev.ptr(out), _dev.ptr(indd), _dev.ptr(fvd), _dev.ptr(xd),
         _dev.stream_handle())
    torch.cuda.synchronize()
    return out.cpu().numpy()


CASES = {
    "basic": dict(),
    "ragged_edges": dict(shape=(37, 131, 10), nv=220),
    "shifted_detector": dict(u_shift=6.0, v_shift=-25.0),
    "flip_u": dict(flip_u=True),
    "coarse_detector": dict(du=1.0, dv=1.0, nu=30, nv=75),
    "mixed_tilted_views": dict(ns=9, tilt_every=4),
    "fine_detector_brick_only": dict(du=0.25, dv=0.25, nu=120, nv=300),
    "nz3": dict(shape=(20, 66, 3), vox=(0.5, 0.5, 4.0)),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_zsep_adjoint_vs_brick_and_fp64(o64, name):
    fv, geo, g2, ind, rng = _case(**CASES[name])
    y = rng.standard_normal(geo[:3]).astype(np.float32)
    y[:, :, 5] = 0.0   # zero rays
    got = _adjoint(fv, geo, g2, ind, y)
    brick = _adjoint(fv, geo, g2, ind, y, brick=(8, 8, 8))
    want = o64.sinobackproject(oracle.LINEAR, 1, y, ind, fv, *geo, *g2)
    assert np.isfinite(got).all()
    assert np.abs(want).max() > 0.1
    eb, e64 = rel_l2(got, brick), rel_l2(got, want)
    print(f"{name}: rel L2 vs brick {eb:.2e}, vs FP64 {e64:.2e}")
    assert eb <= 1e-6, eb
    assert e64 <= 2e-5, e64


@pytest.mark.parametrize("view0,nviews,vstride", [(1, 0, 2), (0, 3, 3), (4, 1, 1)])
def test_zsep_adjoint_view_subsets(o64, view0, nviews, vstride):
    fv, geo, g2, ind, rng = _case(ns=8, tilt_every=5)
    y = rng.standard_normal(geo[:3]).astype(np.float32)
    got = _adjoint(fv, geo, g2, ind, y, view0=view0, nviews=nviews, vstride=vstride)
    n = nviews if nviews > 0 else len(range(view0, geo[0], vstride))
    views = [view0 + i * vstride for i in range(n)]
    keep = np.zeros(geo[0], bool)
    keep[views] = True
    yy = np.where(keep[:, None, None], y, 0).astype(np.float32)
    want = o64.sinobackproject(oracle.LINEAR, 1, yy, ind, fv, *geo, *g2)
    e = rel_l2(got, want)
    print(f"views {views}: rel L2 vs FP64 {e:.2e}")
    assert e <= 2e-5, e


def test_zsep_adjoint_no_indicator_and_unaligned_output():
    fv, geo, g2, ind, rng = _case(shape=(24, 70, 12))
    y = rng.standard_normal(geo[:3]).astype(np.float32)
    got = _adjoint(fv, geo, g2, None, y, offset=1)   # 4-byte aligned only: scalar stores
    aligned = _adjoint(fv, geo, g2, None, y)
    brick = _adjoint(fv, geo, g2, None, y, brick=(8, 8, 8))
    assert np.array_equal(got, aligned)
    assert rel_l2(got, brick) <= 1e-6


@pytest.mark.parametrize("seed", [0, 1])
def test_zsep_dot_product(seed):
    fv, geo, g2, ind, rng = _case(shape=(33, 75, 11), seed=seed)
    nx, ny, nz = g2[6:]
    x = rng.standard_normal((nx, ny, nz)).astype(np.float32)
    y = rng.standard_normal(geo[:3]).astype(np.float32)
    Ax = _forward(fv, geo, g2, ind, x)
    ATy = _adjoint(fv, geo, g2, ind, y)
    lhs = float(np.dot(Ax.astype(np.float64).ravel(), y.astype(np.float64).ravel()))
    rhs = float(np.dot(x.astype(np.float64).ravel(), ATy.astype(np.float64).ravel()))
    scale = float(np.dot(_forward(fv, geo, g2, ind, np.abs(x)).astype(np.float64).ravel(),
                         np.abs(y).astype(np.float64).ravel()))
    gap = abs(lhs - rhs) / scale
    print(f"seed {seed}: <Ax,y> {lhs:.6e}, <x,ATy> {rhs:.6e}, gap {gap:.2e} of <A|x|,|y|>")
    assert gap <= 1e-6, gap


def test_zsep_workspace_groups():
    """ct2_backproject_work_bytes gives the route's workspace (a 256-byte
    header, then per view the planes' maps, the columns' and rows' set-ups
    and trav * datum of every ray, in 16-byte aligned parts); through
    ct2_backproject_ws a workspace of one view's record runs the views one
    group at a time (the first stores, the rest add) and matches the
    allocating entry; a smaller one is rejected."""
    from ct2 import _dev
    from ct2._lib import CT2Error, call, lib
    fv, geo, g2, ind, rng = _case(ns=6)
    ns, nu, nv = geo[:3]
    nx, ny, nz = g2[6:]
    y = rng.standard_normal(geo[:3]).astype(np.float32)
    d = _desc(geo, g2)
    r4 = lambda n: (n + 3) // 4 * 4   # noqa: E731
    rec = r4(r4(4 * nz + 2 * nu + 2 * nv) + nu * nv) * 4
    assert lib().ct2_backproject_work_bytes(C.byref(d)) == 256 + ns * rec
    want = _adjoint(fv, geo, g2, ind, y)
    yd, fvd, indd = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (y, fv, ind))
    out = torch.full((nx, ny, nz), float("nan"), dtype=torch.float32, device="cuda")
    work = torch.empty((256 + rec) // 4, dtype=torch.float32, device="cuda")
    call("ct2_backproject_ws", C.byref(d), _dev.ptr(out), _dev.ptr(yd), _dev.ptr(indd),
         _dev.ptr(fvd), _dev.ptr(work), 256 + rec, _dev.stream_handle())
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    e = rel_l2(got, want)
    print(f"one-view groups vs one group: rel L2 {e:.2e}")
    assert e <= 1e-6, e
    with pytest.raises(CT2Error):
        call("ct2_backproject_ws", C.byref(d), _dev.ptr(out), _dev.ptr(yd), _dev.ptr(indd),
             _dev.ptr(fvd), _dev.ptr(work), 256 + rec - 16, _dev.stream_handle())
