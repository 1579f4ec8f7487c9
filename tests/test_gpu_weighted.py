"""Weighted ray-bunch and derivative measurements and their ART on the GPU
against the FP32 oracle (itself pinned by the reference tests' own checks:
scipy correlation, row norms, Kaczmarz identities).

Bars: measurements bit-identical (EXACT fine projection, correlation in
kernel order); put_weights' image bit-identical (the level schedule keeps
every overlapping pair of bins in sequential order); weights <= 1e-6
(FP64 sums in another order); ART: against the FP64 oracle, no worse than
twice the sequential FP32 oracle's own error, or 1e-5 (the ray sums reduce
in another order, as for single-ray ART; derivative rows with nearest
sampling cancel almost exactly, so both FP32 paths sit near 1e-3 there)."""
import numpy as np
import pytest

import oracle
from cases import cone_case, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def o32():
    return oracle.restated_raw("f32")


@pytest.fixture(scope="module")
def o64():
    return oracle.restated_raw("f64")


KERNELS = [
    (np.ones((1, 4, 4), np.float32) / 16.0, (1, 4, 4)),
    (np.array([[[-1.0], [1.0]]], np.float32), (1, 1, 1)),
    (np.arange(1, 19, dtype=np.float32).reshape(2, 3, 3) / 7.0, (2, 2, 3)),
]


@pytest.mark.parametrize("interp", ["linear", "nearest"])
@pytest.mark.parametrize("k", range(len(KERNELS)))
def test_stencil_against_oracle(o32, o64, interp, k):
    from ct2 import weighted
    w, strides = KERNELS[k]
    ind, fv, geo, vol, g2 = cone_case(n=24, ns=6, nu=40, nv=36, holes=True)
    ns, nu, nv = geo[:3]
    ip = 0 if interp == "linear" else 1
    small = weighted.stencil_project(vol, w, *strides, fv, *geo, g2, interp=interp)
    proj = o32.sinoproj(ip, 0, None, fv, *geo, vol, *g2)
    want = o32.stencil_correlate(proj, w, *strides)
    assert np.array_equal(small, want)
    wts, img = weighted.stencil_weights(w, *strides, ind, fv, *geo, g2, interp=interp)
    owts, oimg = o32.stencil_weights(ip, 0, w, *strides, ind, fv, *geo, *g2)
    assert np.array_equal(img, oimg), rel_l2(img, oimg)
    np.testing.assert_allclose(wts, owts, rtol=1e-6, atol=0)
    assert (wts > 0).sum() > 0.5 * wts.size
    x = np.zeros(vol.shape, np.float32)
    ox = x.copy()
    x64 = x.astype(np.float64)
    w64 = w.astype(np.float64)
    small64 = o64.stencil_correlate(o64.sinoproj(ip, 0, None, fv, *geo, vol.astype(np.float64),
                                                 *g2), w64, *strides)
    wts64, _ = o64.stencil_weights(ip, 0, w64, *strides, ind, fv, *geo, *g2)
    for _ in range(2):
        x = weighted.stencil_art_step(x, w, *strides, small, wts, ind, fv, *geo, g2,
                                      reg_fac=0.8, interp=interp)
        ox = o32.stencil_art(ip, 0, 0.8, w, *strides, want, owts, ind, fv, *geo, ox, *g2)
        x64 = o64.stencil_art(ip, 0, 0.8, w64, *strides, small64, wts64, ind, fv, *geo, x64, *g2)
        e_gpu, e_o32 = rel_l2(x, x64), rel_l2(ox, x64)
        assert e_gpu <= max(2 * e_o32, 1e-5), (e_gpu, e_o32)


def test_parallel_levels_large_detector(o32):
    """A detector wide enough that many bins share a level: the image of the
    used rows must still be bit-identical to the sequential oracle."""
    from ct2 import weighted
    w, strides = KERNELS[0]
    ind, fv, geo, vol, g2 = cone_case(n=32, ns=3, nu=256, nv=192, du=0.4, dv=0.4, holes=True)
    wts, img = weighted.stencil_weights(w, *strides, ind, fv, *geo, g2)
    owts, oimg = o32.stencil_weights(0, 0, w, *strides, ind, fv, *geo, *g2)
    assert np.array_equal(img, oimg)
    np.testing.assert_allclose(wts, owts, rtol=1e-6, atol=0)


def test_reference_method_surface():
    """The reference tests' calls on tomo3D objects, with their checks in
    FP64: <row, phantom> = smallsino, |row|^2 = weights, and one ART step from
    zero with one used bin is row * smallsino / weights
    (tests/cweighted_art_test.py:100-140, tests/cderiv_art_test.py:106-140)."""
    from ct2 import phantoms
    from ct2.tomo3D import image3D, sinogram3D
    img = image3D((32, 32, 32), 20.0, 20.0, 20.0, -10.0, -10.0, -10.0)
    img.mat = phantoms.shepp_logan(32, 9.0)
    data = sinogram3D("circular_conebeam", {"source_to_detector": 30., "radius": 20.},
                      (16, 32, 32), 2 * np.pi, 20.0, 20.0, 0.0, -10.0, -10.0, 0)
    rw = np.ones((1, 4, 4), np.float32) / 16.
    img.cweighted_project_to(data, rw, 1, 4, 4, 1, 8, 8, 16)
    assert data.smallsino.shape == (16, 8, 8)
    for (I, J, K) in ((3, 2, 5), (11, 6, 1)):
        work = img.duplicate()
        data.indicator = np.zeros((16, 32, 32), np.int8)
        data.indicator[I, 4 * J, 4 * K] = 1
        work.put_weights(data, rw, 1, 4, 4, 1, 1, 8, 8)
        row = work.mat.astype(np.float64)
        assert float((row * img.mat).sum()) == pytest.approx(data.smallsino[I, J, K], rel=2e-5)
        assert float((row * row).sum()) == pytest.approx(data.weights[I, J, K], rel=1e-5)
        step = img.duplicate()
        step.mat = np.zeros_like(img.mat)
        step.c_ss_weighted_art_step(data, rw, 1, 4, 4, 1, 1, 8, 8, reg_fac=1.0)
        np.testing.assert_allclose(step.mat, row * data.smallsino[I, J, K] / data.weights[I, J, K],
                                   rtol=1e-5, atol=1e-7 * np.abs(step.mat).max())
    # derivative data (every ray projected)
    data.indicator = np.ones((16, 32, 32), np.int8)
    img.project_to(data)
    p = data.mat.copy()
    data.mat[:, :31, :] = p[:, 1:, :] - p[:, :31, :]
    data.indicator = np.zeros((16, 32, 32), np.int8)
    data.indicator[5, 10, 12] = 1
    work = img.duplicate()
    work.get_deriv_weights(data, 1, 1, 2, 2)
    row = work.mat.astype(np.float64)
    assert float((row * img.mat).sum()) == pytest.approx(data.mat[5, 10, 12], rel=2e-5)
    assert float((row * row).sum()) == pytest.approx(data.weights[5, 10, 12], rel=1e-5)
    step = img.duplicate()
    step.mat = np.zeros_like(img.mat)
    step.cderiv_art_step(data, 1, 1, 4, 4)
    np.testing.assert_allclose(step.mat, row * data.mat[5, 10, 12] / data.weights[5, 10, 12],
                               rtol=1e-5, atol=1e-7 * np.abs(step.mat).max())
