"""The column-separable route of the fast forward projector (views whose
detector rows run along z) against the per-ray march of k_project_views, and
both against the FP64 oracle (view_proj3D.f:703-910).

For such views every ray of a detector column shares its driving axis and
its b coordinate fma(l, dub, ub0) on every plane, so the route lerps the two
b rows of a plane once per column and lets each ray lerp that row along z.
The samples, their arithmetic and the order of the compensated sums are the
per-ray march's, so the two routes must agree bit for bit, for every storage
type (float32, float16, bfloat16, f16 pairs).  CT2_FWD_SEP=0 selects the
per-ray march.

Cases: detector rows per column from 44 (one row per lane) to 600 (two row
chunks), rows stored in reverse (e_v = -z), a shifted detector whose columns
and rows miss the volume, indicator holes, view subsets, views whose
detector is tilted (the route's wavefronts march those ray by ray), z sizes
that the route does not take (nz % 4 != 0, down to nz = 1) and a volume whose
base is not aligned for its quad loads (both stay on the per-ray march, whose
bits the route must equal), a volume whose edge cuts the four-column
workgroups, and the 256^3 x 512^2 scan at its full size on a few views.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest
import torch

import oracle
from cases import rel_l2
from ct2 import configs3D, phantoms
from ct2.operator import to_storage

pytestmark = pytest.mark.gpu


def _case(shape=(36, 28, 72), ns=9, nu=50, nv=44, du=1.6, dv=1.6, u_shift=0.0, v_shift=0.0,
          tilt_every=0, flip_v=False, seed=0, holes=True, R=300.0, sdd=600.0, vox=1.0):
    fv = configs3D.frame_vectors("circular_conebeam", {"radius": R, "source_to_detector": sdd},
                                 ns, 2 * np.pi, 0.1, 0).astype(np.float64)
    if flip_v:
        fv[:, 12:15] *= -1.0
    if tilt_every:
        for p in range(0, ns, tilt_every):
            eu, ev = fv[p, 9:12].copy(), fv[p, 12:15].copy()
            a = np.deg2rad(2.0)
            fv[p, 12:15] = np.cos(a) * ev + np.sin(a) * eu
            fv[p, 9:12] = np.cos(a) * eu - np.sin(a) * ev
    fv = fv.astype(np.float32)
    nx, ny, nz = shape
    geo = (ns, nu, nv, du, dv, -nu * du / 2 + u_shift, -nv * dv / 2 + v_shift)
    g2 = (vox, vox, vox, -nx * vox / 2, -ny * vox / 2, -nz * vox / 2, nx, ny, nz)
    rng = np.random.default_rng(seed)
    ind = (rng.random((ns, nu, nv)) > (0.05 if holes else -1)).astype(np.int8)
    vol = phantoms.random_ellipsoids(max(shape), max(shape) * vox / 2, 10, seed)
    vol = np.ascontiguousarray(vol[:nx, :ny, :nz]).astype(np.float32)
    return fv, geo, g2, ind, vol


def _project(fv, geo, g2, ind, vold, dtype_code, sep, view0=0, nviews=0, vstride=1):
    from ct2 import _dev
    from ct2._lib import CT2_ARITH_FAST, call, ct2_desc
    ns, nu, nv, du, dv, u0, v0 = geo
    dx, dy, dz, x0, y0, z0, nx, ny, nz = g2
    d = ct2_desc()
    call("ct2_desc_init", C.byref(d), ns, nu, nv, du, dv, u0, v0, nx, ny, nz, dx, dy, dz, x0,
         y0, z0)
    d.arith = CT2_ARITH_FAST
    d.vol_dtype = dtype_code
    d.view0, d.nviews, d.vstride = view0, nviews, vstride
    fvd = torch.from_numpy(fv).cuda()
    indd = None if ind is None else torch.from_numpy(ind).cuda()
    out = torch.full((ns, nu, nv), float("nan"), dtype=torch.float32, device="cuda")
    old = os.environ.get("CT2_FWD_SEP")
    os.environ["CT2_FWD_SEP"] = "1" if sep else "0"
    try:
        call("ct2_project", C.byref(d), _dev.ptr(out), None if indd is None else _dev.ptr(indd),
             _dev.ptr(fvd), _dev.ptr(vold), _dev.stream_handle())
        torch.cuda.synchronize()
    finally:
        if old is None:
            del os.environ["CT2_FWD_SEP"]
        else:
            os.environ["CT2_FWD_SEP"] = old
    return out.cpu().numpy()


def _storages(vol):
    from ct2._lib import CT2_BF16, CT2_F16, CT2_F16Z2, CT2_F32
    v32 = torch.from_numpy(vol).cuda()
    return {
        "f32": (v32, CT2_F32),
        "f16": (to_storage(v32, torch.float16, "nearest"), CT2_F16),
        "bf16": (to_storage(v32, torch.bfloat16, "nearest"), CT2_BF16),
        "f16pairs": (to_storage(v32, torch.float16, "nearest", pairs=True), CT2_F16Z2),
    }


CASES = {
    "rows44": dict(),
    "rows90_reversed": dict(shape=(24, 32, 64), nv=90, flip_v=True),
    "rows170_nz256": dict(shape=(16, 12, 256), nu=30, nv=170, R=500.0, sdd=1000.0),
    "rows600": dict(shape=(20, 20, 48), nu=9, nv=600, dv=0.1),
    "shifted_detector": dict(u_shift=14.0, v_shift=-20.0),
    "tilted_views": dict(ns=10, tilt_every=3),
    "nz70_not_taken": dict(shape=(30, 26, 70)),
    "nz255": dict(shape=(9, 7, 255), nu=20, nv=180, R=500.0, sdd=1000.0),
    "nz1": dict(shape=(14, 12, 1), nv=6),
    "nz3": dict(shape=(14, 12, 3), nv=8, dv=0.8),
    "nz5": dict(shape=(13, 11, 5), nv=10, dv=0.8),
    "no_indicator_edge_cut": dict(shape=(33, 27, 40), nu=47, holes=False),
}


@pytest.mark.parametrize("name", list(CASES))
def test_separable_forward_bit_identical(name):
    fv, geo, g2, ind, vol = _case(**CASES[name])
    ind_use = None if name == "no_indicator_edge_cut" else ind
    for label, (vold, code) in _storages(vol).items():
        a = _project(fv, geo, g2, ind_use, vold, code, sep=True)
        b = _project(fv, geo, g2, ind_use, vold, code, sep=False)
        assert np.all(np.isfinite(a)), (name, label)
        diff = np.flatnonzero(a.view(np.int32) != b.view(np.int32))
        # -0.0 and +0.0 are the same projection
        diff = diff[(a.ravel()[diff] != 0) | (b.ravel()[diff] != 0)]
        assert diff.size == 0, (name, label, diff[:5], a.ravel()[diff[:5]], b.ravel()[diff[:5]])


def test_separable_forward_view_subsets():
    fv, geo, g2, ind, vol = _case(ns=12)
    vold, code = _storages(vol)["f32"]
    for view0, nviews, vstride in ((3, 4, 2), (1, 0, 3), (11, 1, 1)):
        a = _project(fv, geo, g2, ind, vold, code, True, view0, nviews, vstride)
        b = _project(fv, geo, g2, ind, vold, code, False, view0, nviews, vstride)
        assert np.array_equal(a, b, equal_nan=True)
        done = set(range(view0, geo[0], vstride))
        if nviews:
            done = set(sorted(done)[:nviews])
        untouched = [p for p in range(geo[0]) if p not in done]
        assert np.isnan(a[untouched]).all()


def test_separable_forward_against_fp64_oracle():
    fv, geo, g2, ind, vol = _case(shape=(32, 32, 48), ns=8, nu=56, nv=64)
    o64 = oracle.restated("f64")
    ref = o64.view_proj3D.sinoproj_linear(ind, fv, *geo, vol, *g2)
    vold, code = _storages(vol)["f32"]
    a = _project(fv, geo, g2, ind, vold, code, True)
    assert rel_l2(a, ref) < 2e-6


def test_separable_forward_full_size_views():
    """256^3 volume (0.8 mm), 512^2 detector (0.8 mm), R = 500, SDD = 1000,
    every 45th of 360 views: the route against the per-ray march, bit for
    bit, in float32 and in f16 pairs."""
    n, ns = 256, 360
    fv = configs3D.frame_vectors("circular_conebeam", {"radius": 500.0, "source_to_detector": 1000.0},
                                 ns, 2 * np.pi, 0.0, 0).astype(np.float32)
    vol = phantoms.random_ellipsoids(n, n * 0.8 / 2, 40, 20260101).astype(np.float32)
    geo = (ns, 512, 512, 0.8, 0.8, -512 * 0.8 / 2, -512 * 0.8 / 2)
    g2 = (0.8, 0.8, 0.8, -n * 0.4, -n * 0.4, -n * 0.4, n, n, n)
    st = _storages(vol)
    for label in ("f32", "f16pairs"):
        vold, code = st[label]
        a = _project(fv, geo, g2, None, vold, code, True, 0, 8, 45)
        b = _project(fv, geo, g2, None, vold, code, False, 0, 8, 45)
        views = np.arange(0, ns, 45)
        assert np.array_equal(a[views], b[views]), label
        assert np.isfinite(a[views]).all() and a[views].max() > 0


def test_separable_forward_unaligned_base():
    """A volume whose base is one element past an aligned address (nz % 4
    == 0): the per-ray march takes it, with the route's sums."""
    fv, geo, g2, ind, vol = _case(shape=(20, 18, 48))
    for label, (vold, code) in _storages(vol).items():
        # one element (one f16 pair) past the tensor's aligned start
        k = 2 if label == "f16pairs" else 1
        shifted = torch.empty(vold.numel() + k, dtype=vold.dtype, device="cuda")
        shifted[k:].copy_(vold.reshape(-1))
        v1 = shifted[k:].view(vold.shape)
        assert v1.data_ptr() % 8 != 0
        a = _project(fv, geo, g2, ind, v1, code, sep=True)
        b = _project(fv, geo, g2, ind, vold, code, sep=False)
        assert np.array_equal(a, b), label
