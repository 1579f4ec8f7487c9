"""The driver compositions of oracle/drivers.py (ASD-POCS of
jawPhantomCCB.py:191-327, EM+TV), pinned by what they must satisfy rather
than by restating them:

* every ASD-POCS record matches quantities computed another way: the data
  distance is the FP64 oracle's residual norm of the same image over the used
  rays, the TV value the FP64 oracle's total_var3d, the image error the
  numpy FP64 distance to the truth;
* the image stays non-negative with a zero two-voxel border, the data
  distance falls on consistent data, and reg_fac follows red_reg^iter
  exactly while gf = gf0 * adist_1 * red_fact^k for some k <= iterations;
* with no TV steps the loop is ART sweeps with edges and positivity (one
  iteration, so reg_fac is still reg_fac0);
* the SART variant (AsdPocs) with no TV steps is one ordered-subset SART
  sweep with edges and positivity, its reg_fac schedule is exact in
  float32, and its data distances fall on consistent data;
* EM+TV started at the truth of consistent data stays there (EM's fixed
  point makes the first image change, and so the TV step size, vanish: a
  step size taken from anything else moves the image), and the TV steps
  lower the TV value below EM's alone.
"""
import numpy as np
import pytest

import oracle
from cases import cone_case, rel_l2
from oracle import drivers as odr


@pytest.fixture(scope="module")
def o32():
    return oracle.restated_raw("f32")


@pytest.fixture(scope="module")
def o64():
    return oracle.restated_raw("f64")


def _case(n=16, ns=10, nu=28, nv=28, holes=True):
    ind, fv, geo, vol, g2 = cone_case(n=n, ns=ns, nu=nu, nv=nv, holes=holes, du=1.2, dv=1.2)
    return ind, fv, geo, np.ascontiguousarray(vol, np.float32), g2


def test_asd_pocs_records_and_schedule(o32, o64):
    ind, fv, geo, vol, g2 = _case()
    b = o32.sinoproj(0, 0, ind, fv, *geo, vol, *g2)
    x0 = np.zeros_like(vol)
    gf0, rf0, red_fact, red_reg = 0.2, 1.0, 0.98, 0.995
    x, rec, (gf, reg_fac, meps) = odr.asd_pocs(o32, x0, b, ind, fv, geo, g2, max_iter=4,
                                               ngradmax=3, gf=gf0, reg_fac=rf0,
                                               red_fact=red_fact, red_reg=red_reg, truth=vol)
    assert [r[0] for r in rec] == [1, 2, 3, 4]
    assert x.min() >= 0.0
    for sl in (np.s_[:2], np.s_[-2:], np.s_[:, :2], np.s_[:, -2:], np.s_[:, :, :2],
               np.s_[:, :, -2:]):
        assert not x[sl].any()
    ddists = [r[1] for r in rec]
    assert all(d1 < d0 for d0, d1 in zip(ddists, ddists[1:]))
    rf = rf0
    for _ in range(4):
        rf *= red_reg
    assert reg_fac == rf
    assert meps == -1.0
    x1, rec1, (gf1, _, _) = odr.asd_pocs(o32, x0, b, ind, fv, geo, g2, max_iter=1, ngradmax=0,
                                         gf=gf0, reg_fac=rf0, truth=vol)
    adist1 = float(np.linalg.norm(x1.astype(np.float64)))   # from zero
    # no TV step: gdist = 0 < conv * adist, so gf keeps gf0 * adist_1
    assert abs(gf1 - gf0 * adist1) <= 1e-6 * gf1
    ks = [gf / (gf0 * adist1 * red_fact ** k) for k in range(5)]
    assert any(abs(r - 1.0) < 1e-3 for r in ks), ks
    # the first record (taken before the TV descent, so of the image after
    # ART) against independent evaluations of that image
    it, ddist, posttv, ierr = rec1[0]
    r64 = o64.sinoproj(0, 0, ind, fv, *geo, x1.astype(np.float64), *g2) - b
    assert abs(ddist - np.sqrt(np.sum(ind * r64 ** 2))) <= 1e-5 * ddist
    assert abs(posttv - o64.total_var3d(0.0, x1.astype(np.float64), *x1.shape)) <= 1e-5 * posttv
    assert abs(ierr - np.linalg.norm(x1.astype(np.float64) - vol)) <= 1e-6 * ierr


def test_asd_pocs_without_tv_is_art(o32):
    ind, fv, geo, vol, g2 = _case(holes=False)
    b = o32.sinoproj(0, 0, None, fv, *geo, vol, *g2)
    x0 = np.full_like(vol, 0.01)
    x, _, _ = odr.asd_pocs(o32, x0, b, None, fv, geo, g2, max_iter=1, nart=2, ngradmax=0,
                           reg_fac=0.7)
    want = o32.artstep(0, 0, 0.7, b, None, fv, *geo, x0, *g2)
    want = o32.artstep(0, 0, 0.7, b, None, fv, *geo, want, *g2)
    odr.zero_edges(want)
    odr.make_positive(want)
    assert np.array_equal(x, want)


def test_em_tv_fixed_point_and_tv_descent(o32):
    ind, fv, geo, vol, g2 = _case(holes=False)
    truth = vol + 0.05
    b = o32.sinoproj(0, 0, None, fv, *geo, truth, *g2)
    x, hist = odr.em_tv(o32, truth, b, None, fv, geo, g2, n_em=3, n_tv=5)
    sens = o32.sinobackproject(0, 0, np.ones_like(b), None, fv, *geo, *g2)
    seen = sens > 1e-3 * sens.max()
    assert hist[0][0] <= 1e-5 * np.linalg.norm(truth)
    assert rel_l2(x[seen], truth[seen]) <= 1e-5
    # from a flat start the TV steps lower the TV value below EM's alone
    x0 = np.full_like(vol, 0.1)
    x_em, _ = odr.em_tv(o32, x0, b, None, fv, geo, g2, n_em=1, n_tv=0)
    x_tv, h = odr.em_tv(o32, x0, b, None, fv, geo, g2, n_em=1, n_tv=5)
    assert h[0][1] > 0
    tv = lambda v: o32.total_var3d(0.0, v, *v.shape)
    assert tv(x_tv) < tv(x_em)
    assert x_tv.min() >= 0.0


def test_asd_pocs_sart_plumbing(o32):
    from oracle import sart as osart
    ind, fv, geo, vol, g2 = _case()
    b = o32.sinoproj(0, 0, ind, fv, *geo, vol, *g2)
    x0 = np.zeros_like(vol)
    x1, rec1, _ = odr.asd_pocs_sart(o32, x0, b, ind, fv, geo, g2, max_iter=1, nsub=3, ngrad=0,
                                    reg_fac=0.9)
    want = osart.sart_sweeps(o32, 0, 0, 0.9, 1, 3, b, ind, fv, *geo, x0, *g2)
    odr.zero_edges(want)
    odr.make_positive(want)
    assert np.array_equal(x1, want)
    x, rec, (gf, reg_fac, meps) = odr.asd_pocs_sart(o32, x0, b, ind, fv, geo, g2, max_iter=4,
                                                    nsub=3, ngrad=3, reg_fac=1.0)
    rf = np.float32(1.0)
    for _ in range(4):
        rf = np.float32(rf * np.float32(0.995))
    assert np.float32(reg_fac) == rf
    d = [r[1] for r in rec]
    assert all(d1 < d0 for d0, d1 in zip(d, d[1:]))
    assert x.min() >= 0.0 and not x[:2].any()
