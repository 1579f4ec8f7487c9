"""Edge cases of the view-indexed projector and its adjoint, in both
arithmetics, against the oracle: a non-cubic volume with a different voxel
size per axis seen by an off-centre detector (rays that miss the volume and
rays that graze it), a single view with one or a few bins, an indicator with
no used ray, and a detector that no ray through the volume reaches.

Bars, as for the centred cases: exact arithmetic bit-identical to the FP32
oracle (forward) or within 1e-6 relative L2 (adjoint, another summation
order); fast arithmetic against the FP64 oracle within the tolerances the
fast path is held to on centred cases."""
import ctypes as C

import numpy as np
import pytest

import oracle
from cases import rel_l2
from ct2 import configs3D

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def o32():
    return oracle.restated("f32")


@pytest.fixture(scope="module")
def o64():
    return oracle.restated("f64")


def aniso_case(ns=8, nu=30, nv=26, du=1.2, dv=1.1, ushift=7.3, vshift=-3.1, seed=5):
    """Non-cubic grid (20 x 28 x 12) with voxels 1.0 x 0.8 x 1.5 and a detector
    moved off the central ray, so that some bins see no voxel."""
    fv = configs3D.frame_vectors("circular_conebeam", {"radius": 60.0, "source_to_detector": 120.0},
                                 ns, 2 * np.pi, 0.3, 0)
    nx, ny, nz = 20, 28, 12
    dx, dy, dz = 1.0, 0.8, 1.5
    rng = np.random.default_rng(seed)
    vol = rng.random((nx, ny, nz)).astype(np.float32)
    g2 = (dx, dy, dz, -nx * dx / 2, -ny * dy / 2, -nz * dz / 2, nx, ny, nz)
    geo = (ns, nu, nv, du, dv, -nu * du / 2 + ushift, -nv * dv / 2 + vshift)
    ind = np.ones((ns, nu, nv), np.int8)
    ind[rng.random((ns, nu, nv)) < 0.05] = 0
    return ind, fv, geo, vol, g2


def tiny_case(nu, nv):
    fv = configs3D.frame_vectors("circular_conebeam", {"radius": 40.0, "source_to_detector": 80.0},
                                 1, 2 * np.pi, 0.7, 0)
    n = 16
    vol = np.random.default_rng(9).random((n, n, n)).astype(np.float32)
    g2 = (1.0, 1.0, 1.0, -n / 2, -n / 2, -n / 2, n, n, n)
    geo = (1, nu, nv, 1.3, 1.3, -nu * 1.3 / 2, -nv * 1.3 / 2)
    ind = np.ones((1, nu, nv), np.int8)
    return ind, fv, geo, vol, g2


def fast_pair(ind, fv, geo, vol, g2, y, vol_dtype="f32", interp=0):
    """Fast-arithmetic forward of vol and adjoint of y through the C ABI."""
    import torch
    from ct2 import _dev
    from ct2._lib import call, ct2_desc, CT2_ARITH_FAST, CT2_F16, CT2_F32
    ns, nu, nv, du, dv, u0, v0 = geo
    dx, dy, dz, x0, y0, z0, nx, ny, nz = g2
    d = ct2_desc()
    call("ct2_desc_init", C.byref(d), ns, nu, nv, du, dv, u0, v0, nx, ny, nz, dx, dy, dz, x0, y0, z0)
    d.arith = CT2_ARITH_FAST
    d.interp = interp
    tdt = {"f32": torch.float32, "f16": torch.float16}[vol_dtype]
    # every device array stays referenced until the synchronisation below
    vd = torch.from_numpy(vol).cuda().to(tdt)
    fvd, indd, yd = (torch.from_numpy(a).cuda() for a in (fv, ind, y))
    Ax = torch.empty((ns, nu, nv), dtype=torch.float32, device="cuda")
    d.vol_dtype = {"f32": CT2_F32, "f16": CT2_F16}[vol_dtype]
    s = _dev.stream_handle()
    call("ct2_project", C.byref(d), _dev.ptr(Ax), _dev.ptr(indd), _dev.ptr(fvd), _dev.ptr(vd), s)
    d.vol_dtype = CT2_F32
    Aty = torch.empty((nx, ny, nz), dtype=torch.float32, device="cuda")
    call("ct2_backproject", C.byref(d), _dev.ptr(Aty), _dev.ptr(yd), _dev.ptr(indd),
         _dev.ptr(fvd), s)
    torch.cuda.synchronize()
    return Ax.cpu().numpy(), Aty.cpu().numpy()


@pytest.mark.parametrize("interp", ["linear", "nearest"])
def test_anisotropic_offcentre_exact(o32, interp):
    from ct2 import view_proj3D as vp
    ind, fv, geo, vol, g2 = aniso_case()
    a = getattr(vp, f"sinoproj_{interp}")(ind, fv, *geo, vol, *g2)
    b = getattr(o32.view_proj3D, f"sinoproj_{interp}")(ind, fv, *geo, vol, *g2)
    # the off-centre detector leaves bins that see no voxel, next to bins that do
    assert (b == 0).sum() > 0.05 * b.size and (b > 1.0).sum() > 0.3 * b.size
    assert np.array_equal(a, b), np.abs(a - b).max()
    y = np.random.default_rng(2).standard_normal(b.shape).astype(np.float32)
    a = getattr(vp, f"sinobackproject_{interp}")(y, ind, fv, *geo, *g2)
    b = getattr(o32.view_proj3D, f"sinobackproject_{interp}")(y, ind, fv, *geo, *g2)
    assert np.abs(b).max() > 0.1
    assert rel_l2(a, b) <= 1e-6


@pytest.mark.parametrize("vol_dtype,fwd_tol", [("f32", 2e-6), ("f16", 1e-3)])
def test_anisotropic_offcentre_fast(o64, vol_dtype, fwd_tol):
    ind, fv, geo, vol, g2 = aniso_case()
    y = np.random.default_rng(4).standard_normal(geo[:3]).astype(np.float32)
    Ax, Aty = fast_pair(ind, fv, geo, vol, g2, y, vol_dtype)
    ref = o64.view_proj3D.sinoproj_linear(ind, fv, *geo, vol.astype(np.float64), *g2)
    e_fwd = rel_l2(Ax, ref)
    refb = o64.view_proj3D.sinobackproject_linear(y.astype(np.float64), ind, fv, *geo, *g2)
    e_bp = rel_l2(Aty, refb)
    print(f"{vol_dtype}: fwd {e_fwd:.2e}  bp {e_bp:.2e}")
    assert e_fwd <= fwd_tol
    assert e_bp <= 2e-5
    # dot-product test of the fast pair (f32 storage: the same matrix)
    if vol_dtype == "f32":
        lhs = float((Ax.astype(np.float64) * y).sum())
        rhs = float((vol.astype(np.float64) * Aty).sum())
        assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), abs(rhs))


@pytest.mark.parametrize("nu,nv", [(1, 1), (3, 2), (1, 70)])
def test_single_view_few_bins(o32, o64, nu, nv):
    from ct2 import view_proj3D as vp
    ind, fv, geo, vol, g2 = tiny_case(nu, nv)
    a = vp.sinoproj_linear(ind, fv, *geo, vol, *g2)
    b = o32.view_proj3D.sinoproj_linear(ind, fv, *geo, vol, *g2)
    assert np.abs(b).max() > 1.0
    assert np.array_equal(a, b)
    y = np.random.default_rng(6).standard_normal(geo[:3]).astype(np.float32)
    a = vp.sinobackproject_linear(y, ind, fv, *geo, *g2)
    b = o32.view_proj3D.sinobackproject_linear(y, ind, fv, *geo, *g2)
    assert rel_l2(a, b) <= 1e-6
    Ax, Aty = fast_pair(ind, fv, geo, vol, g2, y)
    assert rel_l2(Ax, o64.view_proj3D.sinoproj_linear(ind, fv, *geo, vol.astype(np.float64),
                                                      *g2)) <= 2e-6
    assert rel_l2(Aty, o64.view_proj3D.sinobackproject_linear(y.astype(np.float64), ind, fv,
                                                              *geo, *g2)) <= 2e-5


def test_no_used_ray(o32):
    """An all-zero indicator: zero projections and a zero backprojection, in
    both arithmetics (the reference zero-fills its outputs)."""
    from ct2 import view_proj3D as vp
    ind, fv, geo, vol, g2 = aniso_case(ushift=0.0, vshift=0.0)
    ind[:] = 0
    y = np.random.default_rng(8).standard_normal(geo[:3]).astype(np.float32)
    assert not vp.sinoproj_linear(ind, fv, *geo, vol, *g2).any()
    assert not vp.sinobackproject_linear(y, ind, fv, *geo, *g2).any()
    assert not vp.sinobackproject_nearest(y, ind, fv, *geo, *g2).any()
    Ax, Aty = fast_pair(ind, fv, geo, vol, g2, y)
    assert not Ax.any() and not Aty.any()


def test_detector_missing_volume(o32):
    """A detector so far off axis that no ray crosses the volume: all zeros,
    as in the oracle."""
    from ct2 import view_proj3D as vp
    ind, fv, geo, vol, g2 = aniso_case(ushift=400.0)
    b = o32.view_proj3D.sinoproj_linear(ind, fv, *geo, vol, *g2)
    assert not b.any()
    assert not vp.sinoproj_linear(ind, fv, *geo, vol, *g2).any()
    y = np.ones(geo[:3], np.float32)
    assert not vp.sinobackproject_linear(y, ind, fv, *geo, *g2).any()
    Ax, Aty = fast_pair(ind, fv, geo, vol, g2, y)
    assert not Ax.any() and not Aty.any()
