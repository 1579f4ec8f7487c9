"""Golden fixtures written by the compiled reference Fortran
(tests/golden/make_golden.py).  CPU: the FP32 oracle reproduces every
fixture bit for bit.  GPU: the HIP path against the same fixtures --
forward bit-exact, adjoint/ART/MLEM within the summation-order tolerance."""
import os

import numpy as np
import pytest

import oracle
from cases import rel_l2

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with np.load(os.path.join(HERE, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def c1():
    return _load("c1_reference.npz")


@pytest.fixture(scope="module")
def small():
    return _load("small_reference.npz")


def _geo(z):
    g = z["geo"]
    return (int(g[0]), int(g[1]), int(g[2]), *[float(v) for v in g[3:]])


def _grid(g):
    return (*[float(v) for v in g[:6]], int(g[6]), int(g[7]), int(g[8]))


@pytest.fixture(scope="module")
def o32():
    return oracle.restated("f32")


def test_c1_oracle(c1, o32):
    geo, g2 = _geo(c1), _grid(c1["grid"])
    vp = o32.view_proj3D
    sino = vp.sinoproj_linear(None, c1["frame_vectors"], *geo, c1["vol"], *g2)
    assert np.array_equal(sino, c1["sino"])
    bp = vp.sinobackproject_linear(c1["sino"], None, c1["frame_vectors"], *geo, *g2)
    assert np.array_equal(bp, c1["bp"])
    x = vp.artstep_linear(0.5, c1["sino"], None, c1["frame_vectors"], *geo,
                          np.zeros_like(c1["vol"]), *g2)
    assert np.array_equal(x, c1["art1"])


@pytest.mark.slow
def test_c1_oracle_art10(c1, o32):
    geo, g2 = _geo(c1), _grid(c1["grid"])
    x = np.zeros_like(c1["vol"])
    for _ in range(10):
        x = o32.view_proj3D.artstep_linear(0.5, c1["sino"], None, c1["frame_vectors"], *geo, x, *g2)
    assert np.array_equal(x, c1["art10"])


def test_small_oracle(small, o32):
    z = small
    geo = (int(z["cone_geo"][0]), int(z["cone_geo"][1]), int(z["cone_geo"][2]),
           *[float(v) for v in z["cone_geo"][3:]])
    g2 = _grid(z["cone_grid"])
    ind, fv, vol = z["cone_ind"], z["cone_fv"], z["cone_vol"]
    vp = o32.view_proj3D
    s = vp.sinoproj_linear(ind, fv, *geo, vol, *g2)
    assert np.array_equal(s, z["cone_proj_linear"])
    assert np.array_equal(vp.sinoproj_nearest(ind, fv, *geo, vol, *g2), z["cone_proj_nearest"])
    assert np.array_equal(vp.sinobackproject_nearest(s, ind, fv, *geo, *g2), z["cone_adj_nearest"])
    assert np.array_equal(vp.pd_backproj_linear(s, ind, fv, *geo, *g2), z["cone_pd_linear"])
    assert np.array_equal(vp.pd_backproj_nearest(s, ind, fv, *geo, *g2), z["cone_pd_nearest"])
    assert np.array_equal(vp.artstep_nearest(0.5, s, ind, fv, *geo, vol * 0, *g2), z["cone_art_nearest"])
    assert np.array_equal(vp.emstep_linear(s, ind, fv, *geo, vol + 1, *g2), z["cone_em_linear"])
    assert np.array_equal(vp.emstep_nearest(s, ind, fv, *geo, vol + 1, *g2), z["cone_em_nearest"])
    tg = z["tomo_geo"]
    geo = (int(tg[0]), int(tg[1]), int(tg[2]), *[float(v) for v in tg[3:]])
    g2 = _grid(z["tomo_grid"])
    ind, fv, vol = z["tomo_ind"], z["tomo_fv"], z["tomo_vol"]
    vpt = o32.view_proj3D_tomo
    s = vpt.sinoprojz(ind, fv, *geo, vol, *g2)
    assert np.array_equal(s, z["tomo_projz"])
    assert np.array_equal(vpt.artstepz(0.7, s, ind, fv, *geo, vol * 0, *g2), z["tomo_artz"])
    assert np.array_equal(vpt.emstep_linearz(s, ind, fv, *geo, vol + 1, *g2), z["tomo_emz"])
    g3 = _grid(z["ray_grid"])
    ind, rv, vol = z["ray_ind"], z["ray_rv"], z["ray_vol"]
    nr = rv.shape[0]
    rp = o32.ray_proj3D
    s = rp.sinoproj_linear(ind, rv, nr, vol, *g3)
    assert np.array_equal(s, z["ray_proj_linear"])
    assert np.array_equal(rp.sinoproj_nearest(ind, rv, nr, vol, *g3), z["ray_proj_nearest"])
    assert np.array_equal(rp.sinobackproject_nearest(s, ind, rv, nr, *g3), z["ray_adj_nearest"])
    assert np.array_equal(rp.artstep_linear(0.5, s, ind, rv, nr, vol * 0, *g3), z["ray_art_linear"])
    assert np.array_equal(rp.emstep_nearest(s, ind, rv, nr, vol + 1, *g3), z["ray_em_nearest"])
    tvv = z["tv_vol"]
    assert o32.totalvar.total_var3d(0.01, tvv, 9, 10, 11) == float(z["tv_value"])
    g, b = o32.totalvar.total_var_grad3d(0.01, tvv, 9, 10, 11, 1)
    assert np.array_equal(g, z["tv_grad"]) and b == float(z["tv_bobo"])


# ------------------------------------------------------------------ GPU ----
@pytest.mark.gpu
def test_c1_gpu(c1):
    from ct2 import view_proj3D as vp
    geo, g2 = _geo(c1), _grid(c1["grid"])
    fv = c1["frame_vectors"]
    sino = vp.sinoproj_linear(None, fv, *geo, c1["vol"], *g2)
    assert np.array_equal(sino, c1["sino"])
    bp = vp.sinobackproject_linear(c1["sino"], None, fv, *geo, *g2)
    assert rel_l2(bp, c1["bp"]) <= 1e-6
    x = np.zeros_like(c1["vol"])
    for it in range(10):
        x = vp.artstep_linear(0.5, c1["sino"], None, fv, *geo, x, *g2)
        if it == 0:
            assert rel_l2(x, c1["art1"]) <= 1e-5
    err = rel_l2(x, c1["art10"])
    print(f"C1 10 ART sweeps: rel L2 vs reference Fortran = {err:.2e}")
    assert err <= 1e-5


@pytest.mark.gpu
def test_small_gpu(small):
    from ct2 import ray_proj3D as rp
    from ct2 import totalvar
    from ct2 import view_proj3D as vp
    from ct2 import view_proj3D_tomo as vpt
    z = small
    cg = z["cone_geo"]
    geo = (int(cg[0]), int(cg[1]), int(cg[2]), *[float(v) for v in cg[3:]])
    g2 = _grid(z["cone_grid"])
    ind, fv, vol = z["cone_ind"], z["cone_fv"], z["cone_vol"]
    s = z["cone_proj_linear"]
    assert np.array_equal(vp.sinoproj_linear(ind, fv, *geo, vol, *g2), s)
    assert np.array_equal(vp.sinoproj_nearest(ind, fv, *geo, vol, *g2), z["cone_proj_nearest"])
    assert rel_l2(vp.sinobackproject_nearest(s, ind, fv, *geo, *g2), z["cone_adj_nearest"]) <= 1e-6
    assert np.array_equal(vp.pd_backproj_linear(s, ind, fv, *geo, *g2), z["cone_pd_linear"])
    assert np.array_equal(vp.pd_backproj_nearest(s, ind, fv, *geo, *g2), z["cone_pd_nearest"])
    assert rel_l2(vp.artstep_nearest(0.5, s, ind, fv, *geo, vol * 0, *g2), z["cone_art_nearest"]) <= 1e-5
    assert rel_l2(vp.emstep_linear(s, ind, fv, *geo, vol + 1, *g2), z["cone_em_linear"]) <= 1e-5
    assert rel_l2(vp.emstep_nearest(s, ind, fv, *geo, vol + 1, *g2), z["cone_em_nearest"]) <= 1e-5
    tg = z["tomo_geo"]
    geo = (int(tg[0]), int(tg[1]), int(tg[2]), *[float(v) for v in tg[3:]])
    g2 = _grid(z["tomo_grid"])
    ind, fv, vol = z["tomo_ind"], z["tomo_fv"], z["tomo_vol"]
    s = z["tomo_projz"]
    assert np.array_equal(vpt.sinoprojz(ind, fv, *geo, vol, *g2), s)
    assert rel_l2(vpt.artstepz(0.7, s, ind, fv, *geo, vol * 0, *g2), z["tomo_artz"]) <= 1e-5
    assert rel_l2(vpt.emstep_linearz(s, ind, fv, *geo, vol + 1, *g2), z["tomo_emz"]) <= 1e-5
    g3 = _grid(z["ray_grid"])
    ind, rv, vol = z["ray_ind"], z["ray_rv"], z["ray_vol"]
    nr = rv.shape[0]
    s = z["ray_proj_linear"]
    assert np.array_equal(rp.sinoproj_linear(ind, rv, nr, vol, *g3), s)
    assert np.array_equal(rp.sinoproj_nearest(ind, rv, nr, vol, *g3), z["ray_proj_nearest"])
    assert rel_l2(rp.sinobackproject_nearest(s, ind, rv, nr, *g3), z["ray_adj_nearest"]) <= 1e-6
    assert rel_l2(rp.artstep_linear(0.5, s, ind, rv, nr, vol * 0, *g3), z["ray_art_linear"]) <= 1e-5
    assert rel_l2(rp.emstep_nearest(s, ind, rv, nr, vol + 1, *g3), z["ray_em_nearest"]) <= 1e-5
    tvv = z["tv_vol"]
    assert abs(totalvar.total_var3d(0.01, tvv, 9, 10, 11) - float(z["tv_value"])) <= 1e-6 * abs(float(z["tv_value"]))
    g, b = totalvar.total_var_grad3d(0.01, tvv, 9, 10, 11, 1)
    assert np.array_equal(g, z["tv_grad"]) and b == float(z["tv_bobo"])
