"""HIP ART / MLEM / voxel-driven backprojection / TV / FDK against the oracle.

Bars (relative L2 unless stated):
* pd_backproj, TV gradient: bit-identical to the FP32 oracle (same per-voxel
  operation order, sequential sums).
* ART: <= 1e-5 after several full sweeps (only the per-ray raysum/total
  summation order differs from the reference; ART is order-exact).
* MLEM: <= 1e-5 (adjoint sums in another order).
* FDK: <= 1e-5 against the float64 numpy restatement.
"""
import numpy as np
import pytest

import oracle
import oracle.fdk as ofdk
from cases import cone_case, ray_case, rel_l2, tomo_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def o32():
    return oracle.restated("f32")


@pytest.mark.parametrize("interp", ["linear", "nearest"])
def test_pd_backproj_bit_exact(o32, interp):
    from ct2 import view_proj3D as vp
    ind, fv, geo, vol, g2 = cone_case()
    sino = o32.view_proj3D.sinoproj_linear(ind, fv, *geo, vol, *g2)
    a = getattr(vp, f"pd_backproj_{interp}")(sino, ind, fv, *geo, *g2)
    b = getattr(o32.view_proj3D, f"pd_backproj_{interp}")(sino, ind, fv, *geo, *g2)
    assert np.abs(b).max() > 1.0
    assert np.array_equal(a, b), rel_l2(a, b)


@pytest.mark.parametrize("interp", ["linear", "nearest"])
def test_art_matches_oracle(o32, interp):
    from ct2 import view_proj3D as vp
    ind, fv, geo, vol, g2 = cone_case(n=24, ns=8, nu=30, nv=28)
    sino = o32.view_proj3D.sinoproj_linear(ind, fv, *geo, vol, *g2)
    x_gpu = np.zeros_like(vol)
    x_ora = np.zeros_like(vol)
    for _ in range(3):
        x_gpu = getattr(vp, f"artstep_{interp}")(0.5, sino, ind, fv, *geo, x_gpu, *g2)
        x_ora = getattr(o32.view_proj3D, f"artstep_{interp}")(0.5, sino, ind, fv, *geo, x_ora, *g2)
    assert np.abs(x_ora).max() > 0.1
    assert rel_l2(x_gpu, x_ora) <= 1e-5


def test_art_single_ray_consistency():
    """tests/art_test.py:58-70: one ray on, relaxation 1 -> re-projection
    of that ray equals its datum."""
    from ct2 import view_proj3D as vp
    ind, fv, geo, vol, g2 = cone_case(n=24, ns=4, nu=20, nv=20, holes=False)
    ns, nu, nv = geo[:3]
    one = np.zeros((ns, nu, nv), np.int8)
    one[2, 9, 11] = 1
    data = np.full((ns, nu, nv), 3.25, np.float32)
    x = vp.artstep_linear(1.0, data, one, fv, *geo, np.zeros_like(vol), *g2)
    p = vp.sinoproj_linear(one, fv, *geo, x, *g2)
    assert abs(p[2, 9, 11] / 3.25 - 1.0) < 1e-5


def test_art_tomo_and_two_volume(o32):
    from ct2 import view_proj3D_tomo as vpt
    ind, fv, geo, vol, g2 = tomo_case(ns=5)
    sino = o32.view_proj3D_tomo.sinoprojz(ind, fv, *geo, vol, *g2)
    a = vpt.artstepz(0.7, sino, ind, fv, *geo, np.zeros_like(vol), *g2)
    b = o32.view_proj3D_tomo.artstepz(0.7, sino, ind, fv, *geo, np.zeros_like(vol), *g2)
    assert rel_l2(a, b) <= 1e-5
    rng = np.random.default_rng(3)
    va = rng.random((20, 16, 4)).astype(np.float32)
    ga = (1.6, 1.6, 3.0, -16.0, -12.8, 16.0, 20, 16, 4)
    ra = vpt.artstepz2vol(0.7, sino, ind, fv, *geo, vol, *g2, va, *ga)
    oa = o32.view_proj3D_tomo.artstepz2vol(0.7, sino, ind, fv, *geo, vol, *g2, va, *ga)
    assert rel_l2(ra[0], oa[0]) <= 1e-5 and rel_l2(ra[1], oa[1]) <= 1e-5


@pytest.mark.parametrize("interp", ["linear", "nearest"])
def test_ray_art(o32, interp):
    from ct2 import ray_proj3D as rp
    ind, rv, nr, vol, g3 = ray_case(nr=800)
    s = o32.ray_proj3D.sinoproj_linear(ind, rv, nr, vol, *g3)
    a = getattr(rp, f"artstep_{interp}")(0.5, s, ind, rv, nr, vol * 0, *g3)
    b = getattr(o32.ray_proj3D, f"artstep_{interp}")(0.5, s, ind, rv, nr, vol * 0, *g3)
    assert rel_l2(a, b) <= 1e-5


@pytest.mark.parametrize("interp", ["linear", "nearest"])
def test_em_matches_oracle(o32, interp):
    from ct2 import view_proj3D as vp
    ind, fv, geo, vol, g2 = cone_case(n=24, ns=8, nu=30, nv=28)
    sino = o32.view_proj3D.sinoproj_linear(ind, fv, *geo, vol, *g2)
    x_gpu = np.ones_like(vol)
    x_ora = np.ones_like(vol)
    for _ in range(3):
        x_gpu = getattr(vp, f"emstep_{interp}")(sino, ind, fv, *geo, x_gpu, *g2)
        x_ora = getattr(o32.view_proj3D, f"emstep_{interp}")(sino, ind, fv, *geo, x_ora, *g2)
    assert rel_l2(x_gpu, x_ora) <= 1e-5


def test_em_tomo_and_raylist(o32):
    from ct2 import ray_proj3D as rp
    from ct2 import view_proj3D_tomo as vpt
    ind, fv, geo, vol, g2 = tomo_case(ns=5)
    sino = o32.view_proj3D_tomo.sinoprojz(ind, fv, *geo, vol, *g2)
    a = vpt.emstep_linearz(sino, ind, fv, *geo, np.ones_like(vol), *g2)
    b = o32.view_proj3D_tomo.emstep_linearz(sino, ind, fv, *geo, np.ones_like(vol), *g2)
    assert rel_l2(a, b) <= 1e-5
    ind, rv, nr, vol, g3 = ray_case(nr=1500)
    s = o32.ray_proj3D.sinoproj_nearest(ind, rv, nr, vol, *g3)
    a = rp.emstep_nearest(s, ind, rv, nr, vol + 1, *g3)
    b = o32.ray_proj3D.emstep_nearest(s, ind, rv, nr, vol + 1, *g3)
    assert rel_l2(a, b) <= 1e-5


def test_em_consistent_data_fixed_point():
    """x with b = Ax is a fixed point of MLEM where the sensitivity is positive."""
    from ct2 import view_proj3D as vp
    ind, fv, geo, vol, g2 = cone_case(n=24, ns=8, nu=40, nv=40, holes=False)
    x = vol + 0.5
    b = vp.sinoproj_linear(None, fv, *geo, x, *g2)
    x1 = vp.emstep_linear(b, None, fv, *geo, x, *g2)
    sens = vp.sinobackproject_linear(np.ones_like(b), None, fv, *geo, *g2)
    m = sens > 1e-3
    assert rel_l2(x1[m], x[m]) <= 1e-5


@pytest.mark.parametrize("lam,nq", [(0.0, 1), (0.01, 0), (0.02, 1)])
def test_tv_grad_bit_exact(o32, lam, nq):
    from ct2 import totalvar
    rng = np.random.default_rng(8)
    vol = rng.random((17, 19, 23)).astype(np.float32)
    g, bobo = totalvar.total_var_grad3d(lam, vol, 17, 19, 23, nq)
    go, bo = o32.totalvar.total_var_grad3d(lam, vol, 17, 19, 23, nq)
    assert bobo == bo
    assert np.array_equal(g, go)


def test_tv_value(o32):
    from ct2 import totalvar
    rng = np.random.default_rng(9)
    vol = rng.random((17, 19, 23)).astype(np.float32)
    for lam in (0.0, 0.01):
        a = totalvar.total_var3d(lam, vol, 17, 19, 23)
        b = o32.totalvar.total_var3d(lam, vol, 17, 19, 23)
        assert abs(a - b) <= 1e-6 * abs(b)


@pytest.mark.parametrize("p", [0.8, 1.0, 1.5])
def test_tv_p_family(o32, p):
    """totalvar.f p-variants on the device against the FP32 oracle (itself
    bit-exact with the compiled Fortran).  powf differs between the device
    and libm by an ulp at most: 1e-6 relative."""
    from ct2 import totalvar
    rng = np.random.default_rng(10)
    n = (17, 19, 23)
    vol = rng.random(n).astype(np.float32) - 0.3
    o = o32.totalvar
    assert np.array_equal(totalvar.gmi(vol, *n), o.gmi(vol, *n))
    for name, extra in (("calc_lp", ()), ("total_var3d_p", ()),
                        ("total_dvar3d_p", (0.8, 0.9, 1.3))):
        a = getattr(totalvar, name)(p, vol, *n, *extra)
        b = getattr(o, name)(p, vol, *n, *extra)
        assert abs(a - b) <= 1e-6 * abs(b), name
    for nq in (0, 1):
        for name, extra in (("lp_grad3d", ()), ("total_p_var_grad3d", ()),
                            ("total_p_dvar_grad3d", (0.8, 0.9, 1.3))):
            g, bobo = getattr(totalvar, name)(p, vol, *n, *extra, nq)
            go, bo = getattr(o, name)(p, vol, *n, *extra, nq)
            assert abs(bobo - bo) <= 1e-6 * abs(bo), name
            assert rel_l2(g, go) <= 1e-6, name


def _fdk_setup(nu=64, nv=48, ns=36):
    du = dv = 0.4
    u00 = -nu * du / 2
    uoff = voff = 0.5
    vu1 = u00 + (np.arange(nu) + uoff) * du
    vv1 = u00 + (np.arange(nv) + voff) * dv      # the reference builds v from u00 (recon3D.py:60)
    vv, vu = np.meshgrid(vv1, vu1)
    rng = np.random.default_rng(12)
    sino = rng.random((ns, nu, nv)).astype(np.float32)
    return du, dv, u00, vu, vv, sino


@pytest.mark.parametrize("dim,window,L", [(1, "", 1), (2, "", 1), (1, "Hann", 1.0), (1, "Hann", 0.5)])
def test_fdk_filter(dim, window, L):
    from ct2 import FDK
    du, dv, u00, vu, vv, sino = _fdk_setup()
    a = FDK.fdk_sino_filter3D(du, dim, sino, window, L)
    b = ofdk.fdk_sino_filter3D(du, dim, sino, window, L)
    assert rel_l2(a, b) <= 1e-5


def test_fdk_filter_underpadded():
    """n = 360 pads to 512 < 2n-1 (the reference's round(), FDK.py:36):
    the circular wrap-around must be reproduced."""
    from ct2 import FDK
    rng = np.random.default_rng(13)
    sino = rng.random((2, 360, 3)).astype(np.float32)
    a = FDK.fdk_sino_filter3D(0.5, 1, sino)
    b = ofdk.fdk_sino_filter3D(0.5, 1, sino)
    assert rel_l2(a, b) <= 1e-5


def test_fdk_weight_and_backprojection():
    from ct2 import FDK
    du, dv, u00, vu, vv, sino = _fdk_setup()
    ns, nu, nv = sino.shape
    R, SDD = 50.0, 100.0
    w = FDK.fdk_sino_weight3D(R, SDD, ns, vu, vv, sino)
    wo = ofdk.fdk_sino_weight3D(R, SDD, ns, vu, vv, sino)
    assert rel_l2(w, wo) <= 1e-6
    nx = ny = 20
    nz = 12
    x00 = y00 = z00 = -6.0
    dx = dy = 0.6
    dz = 1.0
    vx1 = x00 + (np.arange(nx) + 0.5) * dx
    vy1 = y00 + (np.arange(ny) + 0.5) * dy
    vz = z00 + (np.arange(nz) + 0.5) * dz
    vxg, vyg = np.meshgrid(vx1, vy1)
    vx, vy = vxg.flatten(), vyg.flatten()
    vs = 0.0 + np.arange(ns) * (2 * np.pi / ns)
    a = FDK.fdk_backproj_linear3D(ns, nu, nv, du, dv, u00, u00, R, SDD, 0.5, 0.5, vs, vx, vy, vz, wo)
    b = ofdk.fdk_backproj_linear3D(ns, nu, nv, du, dv, u00, u00, R, SDD, 0.5, 0.5, vs, vx, vy, vz, wo)
    assert a.shape == b.shape
    assert rel_l2(a, b) <= 1e-5
