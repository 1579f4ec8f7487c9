"""sinogram3D call-site methods without a GPU (emil_code/tomo3D.py:213-366):
flip_detector, scramble, clear_indicator / config_indicator and the noise
family."""
import numpy as np
import pytest

from ct2 import tomo3D


def _sino(shape=(10, 8, 6), **kw):
    args = dict(config_name="circular_conebeam",
                parms={"radius": 500.0, "source_to_detector": 1000.0}, shape=shape,
                slen=5.0, ulen=2.0, vlen=3.0, s0=1.0, u0=-1.0, v0=-1.5, s_include_endpoints=0)
    args.update(kw)
    return tomo3D.sinogram3D(**args)


def test_flip_detector_swaps_axes_and_restores():
    s = _sino()
    s.mat = np.random.default_rng(0).random(s.mat.shape).astype(np.float32)
    s.indicator[1, 2, 3] = 0
    fv, mat, ind = s.frame_vectors.copy(), s.mat.copy(), s.indicator.copy()
    s.flip_detector()
    assert (s.nu, s.nv) == (6, 8) and s.mat.shape == (10, 6, 8)
    assert np.array_equal(s.frame_vectors[:, 9:12], fv[:, 12:15])
    assert np.array_equal(s.frame_vectors[:, 12:15], fv[:, 9:12])
    assert np.array_equal(s.frame_vectors[:, :9], fv[:, :9])
    assert np.array_equal(s.mat, mat.transpose(0, 2, 1)) and s.indicator[1, 3, 2] == 0
    assert (s.u0, s.v0, s.du, s.dv, s.ulen, s.vlen) == (-1.5, -1.0, 0.5, 0.25, 3.0, 2.0)
    s.flip_detector()
    assert np.array_equal(s.frame_vectors, fv) and np.array_equal(s.mat, mat)
    assert np.array_equal(s.indicator, ind) and (s.nu, s.nv, s.u0, s.du) == (8, 6, -1.0, 0.25)


def test_scramble_interleaves_views():
    s = _sino(shape=(12, 4, 4))
    fv = s.frame_vectors.copy()
    s.scramble(3)
    assert np.array_equal(s.frame_vectors, fv[[0, 3, 6, 9, 1, 4, 7, 10, 2, 5, 8, 11]])
    with pytest.raises(ValueError):
        _sino(shape=(10, 4, 4)).scramble(3)


def test_config_indicator_ranges():
    # ds = 0.5, du = 0.25, dv = 0.5: grid values are exact in binary
    s = _sino()
    s.clear_indicator()
    assert s.indicator.sum() == 0 and s.indicator.dtype == np.int8
    # s in [2, 3] -> views 2..4; u in [-0.5, 0.25] -> columns 2..5; v in [-1, 0.5)
    # -> rows 1..3 (the v range is open at the top)
    s.config_indicator(2.0, 3.0, -0.5, 0.25, -1.0, 0.5, ray_op='replace')
    want = np.zeros(s.mat.shape, bool)
    want[2:5, 2:6, 1:4] = True
    assert np.array_equal(s.indicator.astype(bool), want)
    # 'add' keeps what is on; snapping to the grid rounds halves away from
    # zero: (1.25 - 1.0)/0.5 = 0.5 -> view 1, which Python 3's round would make 0
    s.config_indicator(1.25, 1.3, -1.0, -1.0, -1.5, -0.9, ray_op='add')
    want[1, 0, 0] = True
    assert np.array_equal(s.indicator.astype(bool), want)
    # ranges beyond the sinogram are clamped to it
    s.config_indicator(-9.0, 99.0, -9.0, 99.0, -9.0, 99.0, ray_op='replace')
    want = np.ones(s.mat.shape, bool)
    want[:, :, -1] = False
    assert np.array_equal(s.indicator.astype(bool), want)


def _loop_noise(kind, mat, ind, arg):
    """The reference's scalar loops (tomo3D.py:334-366) on numpy's global
    generator, with the numpy 1.x promotion the reference ran under: a
    float32 element times or plus a Python float is computed in float64
    (numpy 2 keeps such scalar arithmetic in float32)."""
    m = mat.copy()
    for i in range(m.shape[0]):
        for j in range(m.shape[1]):
            for k in range(m.shape[2]):
                if kind == "mult":
                    m[i, j, k] = float(m[i, j, k]) * max(0., np.random.normal(1., arg))
                elif ind[i, j, k]:
                    std = arg if kind == "abs" else arg * (1.0 + float(m[i, j, k]) / 2.)
                    m[i, j, k] = float(m[i, j, k]) + np.random.normal(0., std)
    return m


@pytest.mark.parametrize("kind,meth,arg", [("mult", "add_noise", 0.4), ("abs", "add_abs_noise", 0.1),
                                           ("rel", "add_rel_noise", 0.02)])
def test_noise_follows_reference_draw_order(kind, meth, arg):
    s = _sino(shape=(3, 5, 4))
    s.mat = np.random.default_rng(1).random(s.mat.shape).astype(np.float32) * 3
    s.indicator = (np.random.default_rng(2).random(s.mat.shape) > 0.3).astype(np.int8)
    np.random.seed(77)
    want = _loop_noise(kind, s.mat, s.indicator, arg)
    np.random.seed(77)
    getattr(s, meth)(arg)
    assert s.mat.dtype == np.float32
    assert np.array_equal(s.mat, want)
    if kind != "mult":
        off = s.indicator == 0
        assert np.array_equal(s.mat[off], _sino(shape=(3, 5, 4)).mat[off] + np.random.default_rng(1)
                              .random(s.mat.shape).astype(np.float32)[off] * 3)
    else:
        assert (s.mat >= 0).all()
