"""Pins the oracle where the compiled reference is not the whole story:
closed forms, invariants and the mathematics the reference relies on.

* Joseph projector, axis-aligned rays through voxel centres: the line sum
  d * sum_i V[i,j,k], with the reference's validity window that drops the
  last voxel row (view_proj3D.f:822-823).
* Adjoint: <Ax,y> = <x,A^T y> for the FP64 oracle.
* ART single-ray invariant (emil_code/tests/art_test.py:58-70) and the
  general relaxation identity a.x_new = lambda*b + (1-lambda)*a.x_old.
* MLEM: consistent data are a fixed point.
* TV gradient: finite differences of total_var3d in FP64
  (emil_code/tests/tpvGrad.py:8-71); a linear ramp has zero interior gradient.
* FDK: the filtered impulse is the Kak-Slaney kernel times du when
  npad >= 2n-1; the full FDK chain reconstructs a ball from its analytic
  cone-beam projections.
"""
import numpy as np
import pytest

import oracle
import oracle.fdk as ofdk
from cases import cone_case, rel_l2


@pytest.fixture(scope="module")
def o64():
    return oracle.restated_raw("f64")


@pytest.fixture(scope="module")
def o32m():
    return oracle.restated("f32")


def test_axis_aligned_line_sums(o64):
    rng = np.random.default_rng(0)
    nx, ny, nz = 7, 6, 5
    # dyadic spacings: every voxel centre is exact in float32, so the
    # rays sit exactly on the centres (an inexact centre just below the
    # first cell boundary would drop that row, view_proj3D.f:822)
    d = (1.5, 0.75, 1.25)
    o = (-3.0, -2.0, 1.0)
    vol = rng.random((nx, ny, nz))
    rays, expect = [], []
    # x-driven rays through the centres of (j,k); y-driven through (i,k)
    for j in range(ny):
        for k in range(nz):
            y = o[1] + (j + 0.5) * d[1]
            z = o[2] + (k + 0.5) * d[2]
            rays.append((-500.0, y, z, 500.0, y, z))
            ok = j < ny - 1 and k < nz - 1
            expect.append(d[0] * vol[:, j, k].sum() if ok else 0.0)
    for i in range(nx):
        for k in range(nz):
            x = o[0] + (i + 0.5) * d[0]
            z = o[2] + (k + 0.5) * d[2]
            rays.append((x, 400.0, z, x, -400.0, z))
            ok = i < nx - 1 and k < nz - 1
            expect.append(d[1] * vol[i, :, k].sum() if ok else 0.0)
    rv = np.array(rays, np.float32)
    # exact float32 centres in the ray list
    got = o64.ray_sinoproj(oracle.LINEAR, None, rv, len(rays), vol, *d, *o, nx, ny, nz)
    assert np.allclose(got, expect, rtol=1e-6, atol=1e-9)


def test_constant_volume_nearest(o64):
    """Nearest sampling of a constant volume: c * trav * (#planes inside)."""
    n = 8
    vol = np.full((n, n, n), 2.0)
    # diagonal-ish ray in the x-y plane through the middle
    rv = np.array([[-100.0, -30.0, 0.1, 100.0, 30.0, 0.1]], np.float32)
    got = o64.ray_sinoproj(oracle.NEAREST, None, rv, 1, vol, 1.0, 1.0, 1.0, -4.0, -4.0, -4.0, n, n, n)
    S = rv[0, :3].astype(np.float64)
    B = rv[0, 3:].astype(np.float64)
    slope = (B[1] - S[1]) / (B[0] - S[0])
    trav = np.sqrt(1 + slope ** 2 + ((B[2] - S[2]) / (B[0] - S[0])) ** 2)
    count = 0
    for i in range(n):
        x = -4.0 + i + 0.5
        y = S[1] + slope * (x - S[0])
        count += 0 <= np.floor(y + 4.0) <= n - 1
    assert got[0] == pytest.approx(2.0 * trav * count, rel=1e-12)


@pytest.mark.parametrize("interp", ["linear", "nearest"])
def test_oracle_dot_product(o64, interp):
    ind, fv, geo, vol, g2 = cone_case(n=16, ns=6, nu=24, nv=20)
    m = oracle.LINEAR if interp == "linear" else oracle.NEAREST
    rng = np.random.default_rng(1)
    x = rng.standard_normal(vol.shape)
    y = rng.standard_normal(geo[:3])
    Ax = o64.sinoproj(m, 0, ind, fv, *geo, x, *g2)
    Aty = o64.sinobackproject(m, 0, y, ind, fv, *geo, *g2)
    lhs, rhs = float((Ax * y).sum()), float((x * Aty).sum())
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_art_single_ray_and_relaxation(o64):
    ind, fv, geo, vol, g2 = cone_case(n=16, ns=3, nu=12, nv=12, holes=False)
    ns, nu, nv = geo[:3]
    one = np.zeros((ns, nu, nv), np.int8)
    one[1, 5, 6] = 1
    b = np.full((ns, nu, nv), 2.5)
    x0 = np.random.default_rng(2).random(vol.shape)
    for lam in (1.0, 0.3):
        x1 = o64.artstep(oracle.LINEAR, 0, lam, b, one, fv, *geo, x0, *g2)
        p0 = o64.sinoproj(oracle.LINEAR, 0, one, fv, *geo, x0, *g2)[1, 5, 6]
        p1 = o64.sinoproj(oracle.LINEAR, 0, one, fv, *geo, x1, *g2)[1, 5, 6]
        assert p1 == pytest.approx(lam * 2.5 + (1 - lam) * p0, rel=1e-12)


def test_em_fixed_point(o64):
    ind, fv, geo, vol, g2 = cone_case(n=16, ns=6, nu=30, nv=30, holes=False)
    x = vol.astype(np.float64) + 0.5
    b = o64.sinoproj(oracle.LINEAR, 0, None, fv, *geo, x, *g2)
    x1 = o64.emstep(oracle.LINEAR, 0, b, None, fv, *geo, x, *g2)
    sens = o64.sinobackproject(oracle.LINEAR, 0, np.ones_like(b), None, fv, *geo, *g2)
    m = sens > 1e-3
    assert np.abs(x1[m] - x[m]).max() <= 1e-10 * np.abs(x).max()


def test_tv_gradient_finite_differences(o64):
    rng = np.random.default_rng(3)
    nx, ny, nz = 3, 4, 5
    x = rng.random((nx, ny, nz))
    lam = 0.0
    g, _ = o64.total_var_grad3d(lam, x, nx, ny, nz, 0, eps=1e-9)
    d = rng.standard_normal(x.shape)
    h = 1e-6
    fp = o64.total_var3d(lam, x + h * d, nx, ny, nz, eps=1e-9)
    fm = o64.total_var3d(lam, x - h * d, nx, ny, nz, eps=1e-9)
    # the f64 facade rounds the value through float like the reference's
    # Real*4 result; use the un-rounded FP64 sum instead
    assert fp != fm
    fd = (_tv64(x + h * d) - _tv64(x - h * d)) / (2 * h)
    assert fd == pytest.approx(float((g * d).sum()), rel=1e-6)


def _tv64(x, eps=1e-9):
    dx = x[1:, 1:, 1:] - x[:-1, 1:, 1:]
    dy = x[1:, 1:, 1:] - x[1:, :-1, 1:]
    dz = x[1:, 1:, 1:] - x[1:, 1:, :-1]
    return float(np.sqrt(eps + dx ** 2 + dy ** 2 + dz ** 2).sum())


def test_tv_ramp_interior_zero(o64):
    n = 6
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    x = (0.3 * i + 0.2 * j - 0.1 * k).astype(np.float64)
    g, _ = o64.total_var_grad3d(0.0, x, n, n, n, 0, eps=1e-9)
    assert np.abs(g[1:-1, 1:-1, 1:-1]).max() < 1e-6


def test_tv_normalised_unit_norm(o32m):
    x = np.random.default_rng(4).random((5, 6, 7)).astype(np.float32)
    g, bobo = o32m.totalvar.total_var_grad3d(0.0, x, 5, 6, 7, 1)
    assert np.linalg.norm(g.astype(np.float64)) == pytest.approx(1.0, rel=1e-6)
    assert bobo > 1.0


def test_fdk_filtered_impulse_is_kak_slaney():
    n, du = 64, 0.5
    s = np.zeros((1, n, 1))
    i0 = 20
    s[0, i0, 0] = 1.0
    y = ofdk.fdk_sino_filter3D(du, 1, s)[0, :, 0]
    m = np.arange(n) - i0
    h = np.where(m == 0, 1 / (4 * du ** 2), np.where(m % 2 != 0, -1 / (np.pi * m * du) ** 2, 0.0))
    assert np.allclose(y, h * du, rtol=1e-10, atol=1e-12)


def test_fdk_reconstructs_ball():
    """Analytic cone-beam projections of a uniform ball -> FDK -> ~1 inside."""
    R, SDD = 100.0, 200.0
    ns, nu, nv, du = 120, 96, 96, 0.5
    rho, rad = 1.0, 8.0
    u00 = -nu * du / 2
    uc = u00 + (np.arange(nu) + 0.5) * du
    vc = u00 + (np.arange(nv) + 0.5) * du
    sino = np.zeros((ns, nu, nv))
    vs = np.arange(ns) * 2 * np.pi / ns
    for p, s in enumerate(vs):
        S = np.array([R * np.cos(s), R * np.sin(s), 0.0])
        D = (R - SDD) * np.array([np.cos(s), np.sin(s), 0.0])
        eu = np.array([-np.sin(s), np.cos(s), 0.0])
        ev = np.array([0.0, 0.0, 1.0])
        B = D[None, None, :] + uc[:, None, None] * eu + vc[None, :, None] * ev
        dirn = B - S
        dirn /= np.linalg.norm(dirn, axis=2, keepdims=True)
        tca = -(S[None, None, :] * dirn).sum(2)           # ball at the origin
        d2 = (S ** 2).sum() - tca ** 2
        sino[p] = rho * 2 * np.sqrt(np.clip(rad ** 2 - d2, 0, None))
    vv, vu = np.meshgrid(vc, uc)
    w = ofdk.fdk_sino_weight3D(R, SDD - R, ns, vu, vv, sino)   # the geometry's own D
    f = ofdk.fdk_sino_filter3D(du, 1, w)
    # detector coordinates scaled to the isocentre: u0, du at the detector
    n = 24
    dx = 0.8
    x1 = -n * dx / 2 + (np.arange(n) + 0.5) * dx
    vxg, vyg = np.meshgrid(x1, x1)
    rec = ofdk.fdk_backproj_linear3D(ns, nu, nv, du, du, u00, u00, R, SDD, 0.5, 0.5, vs,
                                     vxg.flatten(), vyg.flatten(), np.array([0.0]), f)
    cube = ofdk.recon_cube(rec, n, n, 1)[:, :, 0]
    # the reference's unscaled detector coordinates reconstruct the ball at
    # magnification-scaled density; normalise by the detector magnification
    centre = cube[n // 2 - 2:n // 2 + 2, n // 2 - 2:n // 2 + 2].mean()
    outside = cube[0:3, 0:3].mean()
    assert abs(outside) < 0.05 * abs(centre)
    # interior is flat (a ball, not a ring or a spike)
    inner = cube[n // 2 - 4:n // 2 + 4, n // 2 - 4:n // 2 + 4]
    assert inner.std() < 0.05 * abs(inner.mean())
    assert centre > 0


def _tvp64(x, p, d=(1.0, 1.0, 1.0), eps=1e-9):
    dx = (x[1:, 1:, 1:] - x[:-1, 1:, 1:]) / d[0]
    dy = (x[1:, 1:, 1:] - x[1:, :-1, 1:]) / d[1]
    dz = (x[1:, 1:, 1:] - x[1:, 1:, :-1]) / d[2]
    return float((np.sqrt(eps + dx ** 2 + dy ** 2 + dz ** 2) ** p).sum())


@pytest.mark.parametrize("p", [0.8, 1.0, 1.5])
def test_p_gradients_finite_differences(o64, p):
    """The three p-gradients of totalvar.f are the derivatives of their
    values (FP64 restatement, central differences on a random direction)."""
    rng = np.random.default_rng(4)
    n = (4, 5, 6)
    x = rng.random(n) + 0.1
    d = rng.standard_normal(n)
    h = 1e-6
    vox = (0.8, 0.9, 1.3)
    cases = [
        (o64.total_p_var_grad3d(p, x, *n, 0)[0], lambda v: _tvp64(v, p)),
        (o64.total_p_dvar_grad3d(p, x, *n, *vox, 0)[0], lambda v: _tvp64(v, p, vox)),
        (o64.lp_grad3d(p, x, *n, 0)[0], lambda v: float((np.sqrt(1e-9 + v ** 2) ** p).sum())),
    ]
    for g, f in cases:
        fd = (f(x + h * d) - f(x - h * d)) / (2 * h)
        assert fd == pytest.approx(float((g * d).sum()), rel=1e-6)
    # the values themselves: the restated sums agree with the closed forms
    assert o64.total_dvar3d_p(p, x, *n, *vox) == pytest.approx(_tvp64(x, p, vox), rel=1e-6)



# ---- analytic phantoms: brute-force chords from the objects' definitions ----
def _inside(ob, P):
    """Indicator of a phantom object at points P (..., 3), from the definitions:
    plane z' >= 0, ellipsoid sum (x'/a)^2 <= 1, cylinder (x'/a)^2 + (y'/b)^2 <= 1,
    cone z' >= 0 and (x'/a)^2 + (y'/b)^2 <= z'^2, in each surface frame."""
    from ct2 import quadric
    if isinstance(ob, quadric.ellipsoid):
        ca, sa, cb, sb = np.cos(ob.alpha), np.sin(ob.alpha), np.cos(ob.beta), np.sin(ob.beta)
        r = P - np.array([ob.x0, ob.y0, ob.z0])
        r1x, r1y = cb * r[..., 0] + sb * r[..., 1], -sb * r[..., 0] + cb * r[..., 1]
        r2x, r2z = ca * r1x + sa * r[..., 2], -sa * r1x + ca * r[..., 2]
        return (r2x / ob.ax) ** 2 + (r1y / ob.ay) ** 2 + (r2z / ob.az) ** 2 <= 1.0
    ok = np.ones(P.shape[:-1], bool)
    for s in ob.surfaces:
        c = np.array([s.xc + ob.x0, s.yc + ob.y0, s.zc + ob.z0])
        f = (P - c) @ np.array([[s.xu1, s.xu2, s.xu3], [s.yu1, s.yu2, s.yu3],
                                [s.zu1, s.zu2, s.zu3]]).T
        if s.label == "plane":
            ok &= f[..., 2] >= 0
        elif s.label == "ellipsoid":
            ok &= (f[..., 0] / s.ax) ** 2 + (f[..., 1] / s.ay) ** 2 + (f[..., 2] / s.az) ** 2 <= 1
        elif s.label == "cylinder":
            ok &= (f[..., 0] / s.ax) ** 2 + (f[..., 1] / s.ay) ** 2 <= 1
        else:
            ok &= (f[..., 2] >= 0) & ((f[..., 0] / s.ax) ** 2 + (f[..., 1] / s.ay) ** 2
                                      <= f[..., 2] ** 2)
    return ok


def test_analytic_chords_brute_force(o64):
    """Every object's projection equals att times the length of the ray inside
    it, measured by dense sampling along the ray (step 0.006, so the chord is
    known to about two steps)."""
    from cases import ellipsoid_args, f2py_surface_args, quadric_objects
    from ct2 import quadric
    ind, fv, geo, _, _ = cone_case(n=32, ns=3, holes=False)
    ns, nu, nv, du, dv, u0, v0 = geo
    fv64 = np.asarray(fv, np.float64)
    j, k = np.meshgrid(np.arange(0, nu, 3), np.arange(0, nv, 3), indexing="ij")
    nsamp = 10000
    for ob in quadric_objects():
        if isinstance(ob, quadric.ellipsoid):
            proj = o64.ellipsoid_proj(ind, fv, nu, nv, du, dv, u0, v0, *ellipsoid_args(ob))
        else:
            proj = o64.gen_shape_proj(ind, fv, nu, nv, du, dv, u0, v0, ob.att,
                                      *f2py_surface_args(ob))
        for p in range(ns):
            S = fv64[p, 0:3]
            u = u0 + (j + 0.5) * du
            v = v0 + (k + 0.5) * dv
            B = fv64[p, 6:9] + u[..., None] * fv64[p, 9:12] + v[..., None] * fv64[p, 12:15]
            d = B - S
            L = np.linalg.norm(d, axis=-1)
            tc = -(S * d).sum(-1) / (L * L)            # closest approach to the origin
            half = 30.0 / L
            t = tc[..., None] + half[..., None] * np.linspace(-1, 1, nsamp)
            P = S + t[..., None] * d[..., None, :]
            step = 2 * 30.0 / (nsamp - 1)
            chord = _inside(ob, P).sum(-1) * step
            np.testing.assert_allclose(proj[p][j, k], ob.att * chord, atol=3 * step * abs(ob.att))


def test_analytic_closed_forms(o64):
    """A ball's chord is 2 sqrt(r^2 - d^2) for a ray at distance d from the
    centre, through ellipsoid_proj and through a general object with one
    ellipsoidal surface; a voxel embedding of an axis-aligned box of planes
    fills exactly the voxels whose centres lie inside."""
    from cases import f2py_surface_args
    from ct2 import quadric
    ind, fv, geo, _, _ = cone_case(n=32, ns=4, holes=False)
    ns, nu, nv, du, dv, u0, v0 = geo
    c, r = np.array([1.5, -2.0, 0.5]), 9.0
    a = o64.ellipsoid_proj(ind, fv, nu, nv, du, dv, u0, v0, r, r, r, *c, 0.7, -0.4, 2.0)
    ball = quadric.gen_object(*c, att=2.0)
    ball.add_surface(quadric.ellipsoidal(0., 0., 0., 0.3, 0.2, 0.1, ax=r, ay=r, az=r))
    b = o64.gen_shape_proj(ind, fv, nu, nv, du, dv, u0, v0, 2.0, *f2py_surface_args(ball))
    fv64 = np.asarray(fv, np.float64)
    j, k = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    for p in range(ns):
        S = fv64[p, 0:3]
        B = (fv64[p, 6:9] + (u0 + (j + 0.5) * du)[..., None] * fv64[p, 9:12]
             + (v0 + (k + 0.5) * dv)[..., None] * fv64[p, 12:15])
        d = (B - S) / np.linalg.norm(B - S, axis=-1)[..., None]
        w = c - S
        dist2 = (w * w).sum(-1) - ((w * d).sum(-1)) ** 2
        want = 2.0 * 2.0 * np.sqrt(np.maximum(r * r - dist2, 0.0))
        np.testing.assert_allclose(a[p], want, atol=1e-7)
        np.testing.assert_allclose(b[p], want, atol=1e-7)
    box = quadric.gen_object(0.3, -0.2, 0.1, att=1.5)
    for (alpha, beta), cc in (((0., 0.), (0, 0, -4.)), ((0., np.pi), (0, 0, 5.)),
                              ((0., np.pi / 2), (-6., 0, 0)), ((0., -np.pi / 2), (3., 0, 0)),
                              ((np.pi / 2, np.pi / 2), (0, -2., 0)),
                              ((np.pi / 2, -np.pi / 2), (0, 7., 0))):
        box.add_surface(quadric.planar(*cc, alpha, beta, 0.))
    g = (24, 22, 20, 0.75, 0.8, 0.9, -9.1, -8.7, -9.05)
    vol = o64.gen_obj_embed(*g, 1.5, *f2py_surface_args(box))
    x = -9.1 + (np.arange(24) + 0.5) * 0.75
    y = -8.7 + (np.arange(22) + 0.5) * 0.8
    z = -9.05 + (np.arange(20) + 0.5) * 0.9
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    want = 1.5 * ((X >= -5.7) & (X <= 3.3) & (Y >= -2.2) & (Y <= 6.8) & (Z >= -3.9) & (Z <= 5.1))
    assert np.array_equal(vol, want)
    e = o64.ellipsoid_embed(*g, 1.5, 0., 0., 0.5, 0.5, 0.5, 5.0, 4.0, 3.0)
    inside = ((X - 0.5) / 5.0) ** 2 + ((Y - 0.5) / 4.0) ** 2 + ((Z - 0.5) / 3.0) ** 2 < 1.0
    assert np.array_equal(e, 1.5 * inside)


def test_raybunch_invariants(o64):
    """sinoscale times the source-to-bin distance (numpy, FP64) gives the
    input back; sinoave of non-overlapping windows is the block mean, and
    sinodist of it spreads each mean over its block, so sinoave(sinodist(y))
    is y."""
    ind, fv, geo, _, _ = cone_case(n=32, ns=6, nu=12, nv=10)
    ns, nu, nv, du, dv, u0, v0 = geo
    rng = np.random.default_rng(8)
    s = rng.random((ns, nu, nv)) + 0.5
    sc = o64.sinoscale(s, fv, ns, nu, nv, du, dv, u0, v0)
    f = np.asarray(fv, np.float64)
    u = u0 + (np.arange(nu) + 0.5) * du
    v = v0 + (np.arange(nv) + 0.5) * dv
    B = (f[:, None, None, 6:9] + u[None, :, None, None] * f[:, None, None, 9:12]
         + v[None, None, :, None] * f[:, None, None, 12:15])
    d = np.linalg.norm(B - f[:, None, None, 0:3], axis=-1)
    np.testing.assert_allclose(sc * d, s, rtol=1e-12)
    n9 = (3, 4, 5, 2, 3, 2, 2, 3, 2)
    a = o64.sinoave(s, ns, nu, nv, *n9)
    np.testing.assert_allclose(a, s.reshape(3, 2, 4, 3, 5, 2).mean(axis=(1, 3, 5)), rtol=1e-12)
    y = rng.random((3, 4, 5))
    spread = o64.sinodist(y, ns, nu, nv, *n9)
    np.testing.assert_array_equal(spread, np.repeat(np.repeat(np.repeat(y, 2, 0), 3, 1), 2, 2))
    np.testing.assert_allclose(o64.sinoave(spread, ns, nu, nv, *n9), y, rtol=1e-12)


# ---- weighted ray-bunch and derivative measurements (stencils) ----
def _stencil_setup(ns=4, nu=14, nv=12, n=12):
    ind, fv, geo, vol, g2 = cone_case(n=n, ns=ns, nu=nu, nv=nv, phantom="random", holes=False,
                                      du=2.0, dv=2.0)
    return ind, fv, geo, vol.astype(np.float64), g2


@pytest.mark.parametrize("w,strides", [
    (np.ones((1, 4, 4)) / 16.0, (1, 4, 4)),          # tests/cweighted_art_test.py:81
    (np.array([[[-1.0], [1.0]]]), (1, 1, 1)),          # the derivative kernel
    (np.arange(1, 19, dtype=float).reshape(2, 3, 3) / 7.0, (2, 2, 3)),
])
def test_stencil_measurements_and_rows(o64, w, strides):
    """Measurements are the reference test's scipy correlation of the
    projection; put_weights' image is the adjoint of the used stencils; each
    weight is the squared norm of its row, and <row, x> is the measurement
    (tests/cweighted_art_test.py:92-120)."""
    from scipy import ndimage
    ind, fv, geo, vol, g2 = _stencil_setup()
    ns, nu, nv = geo[:3]
    nss, nsu, nsv = strides
    proj = o64.sinoproj(0, 0, None, fv, *geo, vol, *g2)
    small = o64.stencil_correlate(proj, w, *strides)
    want = ndimage.convolve(proj, w[::-1, ::-1, ::-1], mode="constant")[::nss, ::nsu, ::nsv]
    np.testing.assert_allclose(small, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    rng = np.random.default_rng(3)
    cind = np.zeros((ns, nu, nv), np.int8)
    used = [(rng.integers(small.shape[0]), rng.integers(small.shape[1]), rng.integers(small.shape[2]))
            for _ in range(3)]
    for (I, J, K) in used:
        cind[I * nss, J * nsu, K * nsv] = 1
    wts, img = o64.stencil_weights(0, 0, w, *strides, cind, fv, *geo, *g2)
    # the fine sinogram holding every used bin's stencil, and its adjoint
    S = np.zeros((ns, nu, nv))
    os_ = [e - 1 - e // 2 for e in w.shape]
    for (I, J, K) in set(used):
        for a in range(w.shape[0]):
            for c in range(w.shape[1]):
                for d in range(w.shape[2]):
                    p, j, k = I * nss + a - os_[0], J * nsu + c - os_[1], K * nsv + d - os_[2]
                    if 0 <= p < ns and 0 <= j < nu and 0 <= k < nv:
                        S[p, j, k] += w[a, c, d]
    np.testing.assert_allclose(img, o64.sinobackproject(0, 0, S, None, fv, *geo, *g2),
                               rtol=1e-12, atol=1e-12 * np.abs(img).max())
    for (I, J, K) in set(used):
        Sb = np.zeros((ns, nu, nv))
        for a in range(w.shape[0]):
            for c in range(w.shape[1]):
                for d in range(w.shape[2]):
                    p, j, k = I * nss + a - os_[0], J * nsu + c - os_[1], K * nsv + d - os_[2]
                    if 0 <= p < ns and 0 <= j < nu and 0 <= k < nv:
                        Sb[p, j, k] += w[a, c, d]
        row = o64.sinobackproject(0, 0, Sb, None, fv, *geo, *g2)
        assert wts[I, J, K] == pytest.approx(float((row * row).sum()), rel=1e-12)
        assert float((row * vol).sum()) == pytest.approx(small[I, J, K], rel=1e-10, abs=1e-12)
    assert np.count_nonzero(wts) <= len(set(used))


def test_stencil_art_kaczmarz(o64):
    """From zero with one used bin, the step is row * m / |row|^2 (the
    reference test's check 4); after a sweep with reg_fac 1 the last bin's
    equation holds exactly, and sweeps reduce the coarse residual."""
    w = np.ones((1, 2, 2)) / 4.0
    strides = (1, 2, 2)
    ind, fv, geo, vol, g2 = _stencil_setup()
    ns, nu, nv = geo[:3]
    nx, ny, nz = g2[6:]
    proj = o64.sinoproj(0, 0, None, fv, *geo, vol, *g2)
    small = o64.stencil_correlate(proj, w, *strides)
    one = np.zeros((ns, nu, nv), np.int8)
    one[2, 6, 4] = 1
    wts, row = o64.stencil_weights(0, 0, w, *strides, one, fv, *geo, *g2)
    x = o64.stencil_art(0, 0, 1.0, w, *strides, small, wts, one, fv, *geo, np.zeros((nx, ny, nz)),
                        *g2)
    np.testing.assert_allclose(x, row * small[2, 3, 2] / wts[2, 3, 2], rtol=1e-12, atol=1e-15)
    allb = np.ones((ns, nu, nv), np.int8)
    wts, _ = o64.stencil_weights(0, 0, w, *strides, allb, fv, *geo, *g2)
    x = np.zeros((nx, ny, nz))
    res = []
    for _ in range(3):
        x = o64.stencil_art(0, 0, 1.0, w, *strides, small, wts, allb, fv, *geo, x, *g2)
        m = o64.stencil_correlate(o64.sinoproj(0, 0, None, fv, *geo, x, *g2), w, *strides)
        res.append(np.linalg.norm(m - small))
        last = tuple(np.array(np.nonzero(wts > 0))[:, -1])
        assert m[last] == pytest.approx(small[last], rel=1e-9)
    assert res[2] < res[1] < res[0] < 0.5 * np.linalg.norm(small)
