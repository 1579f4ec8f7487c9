"""The point tests ellipse_voxval and gen_obj_voxval (tomo3DfortranAux.f:179-210,
:281-416) of the oracle: equal to the compiled reference Fortran at random
points near and across the surfaces, and, for surfaces with zero angles
(framer then only shifts by the centre), equal to gen_obj_embed's values at
the voxel centres, which take the frame from unit vectors instead."""
import numpy as np
import pytest

import oracle


def surfaces_args(rng, with_angles=True):
    """A general object in the f2py layout of gen_obj_voxval: two planes, one
    ellipsoid, one cylinder and one cone, each count and semi-axis array with
    its leading dummy entry."""
    n = 5
    ang = rng.uniform(-1.2, 1.2, (3, n)) if with_angles else np.zeros((3, n))
    cen = rng.uniform(-0.6, 0.6, (3, n))
    cen[2, 0], cen[2, 1] = -2.5, -2.0          # planes below the points' range
    ax_el, ay_el, az_el = [1., 3.0], [1., 2.5], [1., 2.8]
    ax_cy, ay_cy = [1., 2.2], [1., 2.6]
    ax_co, ay_co = [1., 0.9], [1., 1.1]
    cen[2, 4] = -3.0                           # cone apex below the points
    return (5, *ang, *cen, 3, 2, ax_el, ay_el, az_el, 2, ax_cy, ay_cy, 2, ax_co, ay_co)


@pytest.fixture(scope="module")
def o32():
    return oracle.restated("f32")


@pytest.mark.skipif(not oracle.ref_available(), reason="compiled reference not built")
def test_voxval_matches_compiled_reference(o32):
    ref = oracle.reference()
    rng = np.random.default_rng(3)
    x, y, z = rng.uniform(-3.5, 3.5, (3, 3000)).astype(np.float32)
    for seed in range(3):
        args = surfaces_args(np.random.default_rng(seed))
        a = o32.tomo3Dfort.gen_obj_voxval(x, y, z, 1.7, *args)
        b = ref.tomo3Dfort.gen_obj_voxval(x, y, z, 1.7, *args)
        assert (b != 0).sum() > 30 and (b == 0).sum() > 30
        assert np.array_equal(a, b)
    e = (1.3, 0.4, -0.7, 0.2, -0.1, 0.3, 2.5, 1.5, 3.0)
    a = o32.tomo3Dfort.ellipse_voxval(x, y, z, *e)
    b = ref.tomo3Dfort.ellipse_voxval(x, y, z, *e)
    assert 0.05 < (b != 0).mean() < 0.95
    assert np.array_equal(a, b)


def test_zero_angle_voxval_equals_embed(o32):
    """With alpha = beta = gam = 0, framer is a shift by the centre, the same
    as the unit-vector frame of gen_obj_embed with identity axes; the ellipse
    point test equals ellipsoid_embed at the voxel centres."""
    rng = np.random.default_rng(11)
    args = surfaces_args(rng, with_angles=False)
    nsurf, cen, tail = args[0], args[4:7], args[7:]
    eye = [np.full(nsurf, v) for v in (1, 0, 0, 0, 1, 0, 0, 0, 1)]
    g = (12, 10, 14, 0.5, 0.55, 0.45, -3.0, -2.75, -3.15)
    vol = o32.tomo3Dfort.gen_obj_embed(*g, 1.7, nsurf, *eye, *cen, *tail)
    i, j, k = np.meshgrid(*(np.arange(n) for n in g[:3]), indexing="ij")
    x = np.float32(g[6]) + (i.astype(np.float32) + np.float32(0.5)) * np.float32(g[3])
    y = np.float32(g[7]) + (j.astype(np.float32) + np.float32(0.5)) * np.float32(g[4])
    z = np.float32(g[8]) + (k.astype(np.float32) + np.float32(0.5)) * np.float32(g[5])
    pts = o32.tomo3Dfort.gen_obj_voxval(x, y, z, 1.7, *args)
    assert 0.05 < (vol != 0).mean() < 0.95
    assert np.array_equal(pts, vol)
    e = (1.3, 0.4, -0.7, 0.2, -0.1, 0.3, 2.5, 1.5, 3.0)
    vol = o32.tomo3Dfort.ellipsoid_embed(*g, *e)
    assert np.array_equal(o32.tomo3Dfort.ellipse_voxval(x, y, z, *e), vol)
