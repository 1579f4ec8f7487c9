"""The point tests ellipse_voxval / gen_obj_voxval and image3D.add_shape on the
GPU against the oracle, bit-exact: random points across angle-framed
surfaces, the same object's add_shape volume against the oracle's point test
at the voxel centres, and an all-outside and an empty point set."""
import numpy as np
import pytest

import oracle
from test_oracle_voxval import surfaces_args

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def o32():
    return oracle.restated("f32")


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_gen_obj_voxval_points(o32, seed):
    from ct2 import tomo3Dfort
    args = surfaces_args(np.random.default_rng(seed))
    x, y, z = np.random.default_rng(10 + seed).uniform(-3.5, 3.5, (3, 20000)).astype(np.float32)
    a = tomo3Dfort.gen_obj_voxval(x, y, z, 1.7, *args)
    b = o32.tomo3Dfort.gen_obj_voxval(x, y, z, 1.7, *args)
    assert (b != 0).sum() > 100 and (b == 0).sum() > 100
    assert np.array_equal(a, b)
    # a scalar point gives a float, as the f2py function does
    assert tomo3Dfort.gen_obj_voxval(float(x[0]), float(y[0]), float(z[0]), 1.7, *args) == b[0]


def test_ellipse_voxval_points(o32):
    from ct2 import tomo3Dfort
    e = (1.3, 0.4, -0.7, 0.2, -0.1, 0.3, 2.5, 1.5, 3.0)
    x, y, z = np.random.default_rng(4).uniform(-3.5, 3.5, (3, 20000)).astype(np.float32)
    a = tomo3Dfort.ellipse_voxval(x, y, z, *e)
    b = o32.tomo3Dfort.ellipse_voxval(x, y, z, *e)
    assert (b != 0).sum() > 100 and (b == 0).sum() > 100
    assert np.array_equal(a, b)
    far = tomo3Dfort.ellipse_voxval(x + 50, y, z, *e)
    assert not far.any()
    assert tomo3Dfort.ellipse_voxval(np.zeros(0), np.zeros(0), np.zeros(0), *e).shape == (0,)


def test_add_shape_matches_point_test(o32):
    """image3D.add_shape of a rotated gen_object and an ellipsoid: the sum of
    the oracle's point tests at the voxel centres, added onto the image."""
    from ct2 import quadric
    from ct2.tomo3D import image3D
    obj = quadric.gen_object(0.2, -0.1, 0.05, 1.5)
    obj.add_surface(quadric.planar(0., 0., -0.6, 0.3, 0.2, 0.1))
    obj.add_surface(quadric.ellipsoidal(0.05, 0., 0., 0.4, -0.3, 0.2, 0.8, 0.6, 0.7))
    obj.add_surface(quadric.cylindrical(0., 0.05, 0., -0.2, 0.5, 0.3, 0.5, 0.55))
    obj.rotate(0.3, 0.2, -0.1)
    ell = quadric.ellipsoid(-0.1, 0.1, 0.0, 0.7, 0.5, 0.4, 0.6, 0.3, -0.2)
    img = image3D(shape=(24, 20, 28), xlen=2.0, ylen=1.8, zlen=2.2)
    base = np.random.default_rng(5).random((24, 20, 28)).astype(np.float32)
    img.mat = base.copy()
    img.add_shape(obj)
    img.add_shape(ell)
    i, j, k = np.meshgrid(*(np.arange(n) for n in (img.nx, img.ny, img.nz)), indexing="ij")
    f = np.float32
    x = f(img.x0) + (i.astype(f) + f(0.5)) * f(img.dx)
    y = f(img.y0) + (j.astype(f) + f(0.5)) * f(img.dy)
    z = f(img.z0) + (k.astype(f) + f(0.5)) * f(img.dz)
    ang = [s for kind in ("plane", "ellipsoid", "cylinder", "cone") for s in obj.surfaces
           if s.label == kind]
    cols = [[getattr(s, a) for s in ang] for a in ("alpha", "beta", "gamma")]
    cen = [[s.xc + obj.x0 for s in ang], [s.yc + obj.y0 for s in ang],
           [s.zc + obj.z0 for s in ang]]
    args = (3, *cols, *cen, 2, 2, [1., 0.8], [1., 0.6], [1., 0.7], 2, [1., 0.5], [1., 0.55],
            1, [1.], [1.])
    want_obj = o32.tomo3Dfort.gen_obj_voxval(x, y, z, 1.5, *args)
    want_ell = o32.tomo3Dfort.ellipse_voxval(x, y, z, 0.7, 0.3, -0.2, -0.1, 0.1, 0.0,
                                             0.5, 0.4, 0.6)
    assert (want_obj != 0).sum() > 200 and (want_ell != 0).sum() > 200
    got = np.asarray(img.mat)
    assert np.array_equal(got, (base + want_obj) + want_ell)
