"""Pins the FP32 oracle to the reference Fortran compiled from its own
sources (oracle/_ref): every function must agree bit for bit."""
import numpy as np
import pytest

import oracle
from cases import cone_case, ray_case, tomo_case

pytestmark = pytest.mark.skipif(not oracle.ref_available(),
                                reason="oracle/_ref not built (reference sources absent)")


@pytest.fixture(scope="module")
def ref():
    return oracle.reference()


@pytest.fixture(scope="module")
def o32():
    return oracle.restated("f32")


@pytest.mark.parametrize("name", ["sinoproj_linear", "sinoproj_nearest"])
def test_view_forward(ref, o32, name):
    ind, fv, geo, vol, g2 = cone_case()
    a = getattr(ref.view_proj3D, name)(ind, fv, *geo, vol, *g2)
    b = getattr(o32.view_proj3D, name)(ind, fv, *geo, vol, *g2)
    assert np.abs(a).max() > 1.0
    assert np.array_equal(a, b)


@pytest.mark.parametrize("name", ["sinobackproject_linear", "sinobackproject_nearest",
                                  "pd_backproj_linear", "pd_backproj_nearest"])
def test_view_backprojections(ref, o32, name):
    ind, fv, geo, vol, g2 = cone_case()
    sino = o32.view_proj3D.sinoproj_linear(ind, fv, *geo, vol, *g2)
    a = getattr(ref.view_proj3D, name)(sino, ind, fv, *geo, *g2)
    b = getattr(o32.view_proj3D, name)(sino, ind, fv, *geo, *g2)
    assert np.abs(a).max() > 1.0
    assert np.array_equal(a, b)


@pytest.mark.parametrize("name", ["artstep_linear", "artstep_nearest"])
def test_view_art(ref, o32, name):
    ind, fv, geo, vol, g2 = cone_case(ns=6)
    sino = o32.view_proj3D.sinoproj_linear(ind, fv, *geo, vol, *g2)
    x0 = np.zeros_like(vol)
    a = getattr(ref.view_proj3D, name)(0.5, sino, ind, fv, *geo, x0, *g2)
    b = getattr(o32.view_proj3D, name)(0.5, sino, ind, fv, *geo, x0, *g2)
    assert np.abs(a).max() > 0.1
    assert np.array_equal(a, b)


@pytest.mark.parametrize("name", ["emstep_linear", "emstep_nearest"])
def test_view_em(ref, o32, name):
    ind, fv, geo, vol, g2 = cone_case(ns=6)
    sino = o32.view_proj3D.sinoproj_linear(ind, fv, *geo, vol, *g2)
    x0 = np.ones_like(vol)
    a = getattr(ref.view_proj3D, name)(sino, ind, fv, *geo, x0, *g2)
    b = getattr(o32.view_proj3D, name)(sino, ind, fv, *geo, x0, *g2)
    assert np.array_equal(a, b)


def test_tomo(ref, o32):
    ind, fv, geo, vol, g2 = tomo_case()
    a = ref.view_proj3D_tomo.sinoprojz(ind, fv, *geo, vol, *g2)
    b = o32.view_proj3D_tomo.sinoprojz(ind, fv, *geo, vol, *g2)
    assert np.abs(a).max() > 1.0 and np.array_equal(a, b)
    x0 = np.zeros_like(vol)
    assert np.array_equal(ref.view_proj3D_tomo.artstepz(0.7, b, ind, fv, *geo, x0, *g2),
                          o32.view_proj3D_tomo.artstepz(0.7, b, ind, fv, *geo, x0, *g2))
    x1 = np.ones_like(vol)
    assert np.array_equal(ref.view_proj3D_tomo.emstep_linearz(b, ind, fv, *geo, x1, *g2),
                          o32.view_proj3D_tomo.emstep_linearz(b, ind, fv, *geo, x1, *g2))
    rng = np.random.default_rng(3)
    va = rng.random((20, 16, 4)).astype(np.float32)
    ga = (1.6, 1.6, 3.0, -16.0, -12.8, 16.0, 20, 16, 4)
    ra = ref.view_proj3D_tomo.artstepz2vol(0.7, b, ind, fv, *geo, vol, *g2, va, *ga)
    oa = o32.view_proj3D_tomo.artstepz2vol(0.7, b, ind, fv, *geo, vol, *g2, va, *ga)
    assert np.array_equal(ra[0], oa[0]) and np.array_equal(ra[1], oa[1])


def test_ray_list(ref, o32):
    ind, rv, nr, vol, g3 = ray_case()
    for name in ("sinoproj_linear", "sinoproj_nearest"):
        a = getattr(ref.ray_proj3D, name)(ind, rv, nr, vol, *g3)
        b = getattr(o32.ray_proj3D, name)(ind, rv, nr, vol, *g3)
        assert np.abs(a).max() > 1.0 and np.array_equal(a, b), name
    s = b
    assert np.array_equal(ref.ray_proj3D.sinobackproject_nearest(s, ind, rv, nr, *g3),
                          o32.ray_proj3D.sinobackproject_nearest(s, ind, rv, nr, *g3))
    for name in ("artstep_linear", "artstep_nearest"):
        a = getattr(ref.ray_proj3D, name)(0.5, s, ind, rv, nr, vol * 0, *g3)
        b = getattr(o32.ray_proj3D, name)(0.5, s, ind, rv, nr, vol * 0, *g3)
        assert np.array_equal(a, b), name
    assert np.array_equal(ref.ray_proj3D.emstep_nearest(s, ind, rv, nr, vol + 1, *g3),
                          o32.ray_proj3D.emstep_nearest(s, ind, rv, nr, vol + 1, *g3))


def test_totalvar(ref, o32):
    rng = np.random.default_rng(5)
    vol = rng.random((9, 11, 13)).astype(np.float32)
    for lam in (0.0, 0.01):
        assert ref.totalvar.total_var3d(lam, vol, 9, 11, 13) == \
            pytest.approx(o32.totalvar.total_var3d(lam, vol, 9, 11, 13), rel=0, abs=0)
        for nq in (0, 1):
            ga, ba = ref.totalvar.total_var_grad3d(lam, vol, 9, 11, 13, nq)
            gb, bb = o32.totalvar.total_var_grad3d(lam, vol, 9, 11, 13, nq)
            assert np.array_equal(ga, gb) and ba == bb


@pytest.mark.parametrize("p", [0.8, 1.0, 1.5])
def test_totalvar_p_family(ref, o32, p):
    """gmi, calc_lp, total_var3d_p, total_dvar3d_p and the three p-gradients
    (totalvar.f:4-472): the FP32 restatement equals the compiled Fortran."""
    rng = np.random.default_rng(6)
    n = (9, 11, 13)
    vol = rng.random(n).astype(np.float32) - 0.3
    a, b = ref.totalvar, o32.totalvar
    assert np.array_equal(a.gmi(vol, *n), b.gmi(vol, *n))
    assert a.calc_lp(p, vol, *n) == b.calc_lp(p, vol, *n)
    assert a.total_var3d_p(p, vol, *n) == b.total_var3d_p(p, vol, *n)
    assert a.total_dvar3d_p(p, vol, *n, 0.8, 0.9, 1.3) == b.total_dvar3d_p(p, vol, *n, 0.8, 0.9, 1.3)
    for nq in (0, 1):
        for name, extra in (("lp_grad3d", ()), ("total_p_var_grad3d", ()),
                            ("total_p_dvar_grad3d", (0.8, 0.9, 1.3))):
            ga, ba = getattr(a, name)(p, vol, *n, *extra, nq)
            gb, bb = getattr(b, name)(p, vol, *n, *extra, nq)
            assert np.array_equal(ga, gb) and ba == bb, name



def test_analytic_phantoms(ref, o32):
    """gen_shape_proj / ellipsoid_proj (view and ray-list forms) and the two
    embedders: the FP32 restatement equals the compiled Fortran bit for bit
    on objects with every surface kind."""
    from cases import ellipsoid_args, f2py_surface_args, quadric_objects
    from ct2 import quadric
    ind, fv, geo, _, _ = cone_case(n=32, ns=5)
    ns, nu, nv, du, dv, u0, v0 = geo
    objs = quadric_objects()
    nonzero = 0
    for ob in objs:
        for p in range(ns):
            args = (ind[p], fv[p], nu, nv, du, dv, u0, v0)
            if isinstance(ob, quadric.ellipsoid):
                a = ref.view_proj3D.ellipsoid_proj(*args, *ellipsoid_args(ob))
                b = o32.view_proj3D.ellipsoid_proj(*args, *ellipsoid_args(ob))
            else:
                a = ref.view_proj3D.gen_shape_proj(*args, ob.att, *f2py_surface_args(ob))
                b = o32.view_proj3D.gen_shape_proj(*args, ob.att, *f2py_surface_args(ob))
            assert np.array_equal(a, b), (type(ob).__name__, p)
            nonzero += int((a != 0).sum())
    assert nonzero > 2000
    rind, rv, nr, _, _ = ray_case(nr=800)
    for ob in objs:
        if isinstance(ob, quadric.ellipsoid):
            a = ref.ray_proj3D.ellipsoid_proj(rind, rv, nr, *ellipsoid_args(ob))
            b = o32.ray_proj3D.ellipsoid_proj(rind, rv, nr, *ellipsoid_args(ob))
        else:
            a = ref.ray_proj3D.gen_shape_proj(rind, rv, nr, ob.att, *f2py_surface_args(ob))
            b = o32.ray_proj3D.gen_shape_proj(rind, rv, nr, ob.att, *f2py_surface_args(ob))
        assert np.array_equal(a, b)
    g = (20, 22, 18, 1.6, 1.5, 1.9, -16., -16.5, -17.)
    for ob in objs:
        if isinstance(ob, quadric.ellipsoid):
            e = ellipsoid_args(ob)
            a = ref.tomo3Dfort.ellipsoid_embed(*g, e[8], e[6], e[7], *e[3:6], *e[0:3])
            b = o32.tomo3Dfort.ellipsoid_embed(*g, e[8], e[6], e[7], *e[3:6], *e[0:3])
        else:
            a = ref.tomo3Dfort.gen_obj_embed(*g, ob.att, *f2py_surface_args(ob))
            b = o32.tomo3Dfort.gen_obj_embed(*g, ob.att, *f2py_surface_args(ob))
        assert np.array_equal(a, b) and np.array_equal(np.signbit(a), np.signbit(b))
        assert (a != 0).any()


@pytest.mark.parametrize("win", [(1, 2, 2, 1, 2, 2), (2, 3, 2, 3, 4, 2), (1, 4, 4, 1, 4, 4),
                                 (2, 2, 3, 1, 1, 1)])
def test_raybunch(ref, o32, win):
    """sinoscale, sinoave and sinodist (tomo3DfortranAux.f:1-176): FP32
    restatement equals the compiled Fortran, overlapping windows included."""
    ind, fv, geo, _, _ = cone_case(n=32, ns=9, nu=20, nv=18)
    ns, nu, nv, du, dv, u0, v0 = geo
    rng = np.random.default_rng(7)
    s = rng.standard_normal((ns, nu, nv)).astype(np.float32)
    assert np.array_equal(ref.tomo3Dfort.sinoscale(s, fv, ns, nu, nv, du, dv, u0, v0),
                          o32.tomo3Dfort.sinoscale(s, fv, ns, nu, nv, du, dv, u0, v0))
    nds, ndu, ndv, nws, nwu, nwv = win
    nss, nus, nvs = (ns - nws) // nds + 1, (nu - nwu) // ndu + 1, (nv - nwv) // ndv + 1
    n9 = (nss, nus, nvs, nds, ndu, ndv, nws, nwu, nwv)
    a = ref.tomo3Dfort.sinoave(s, ns, nu, nv, *n9)
    assert np.array_equal(a, o32.tomo3Dfort.sinoave(s, ns, nu, nv, *n9))
    assert np.array_equal(ref.tomo3Dfort.sinodist(a, ns, nu, nv, *n9),
                          o32.tomo3Dfort.sinodist(a, ns, nu, nv, *n9))
