"""Flipped-detector symmetry on the GPU, the check of the reference's
tests/clin_backproject_test.py:106-117: a scan and the same scan with the
detector's u and v exchanged (sinogram3D.flip_detector, tomo3D.py:213-238,
or the circular_conebeam_flipped geometry, configs3D.py:196-234) hold the
same rays, so projections come out transposed and backprojections equal."""
import numpy as np
import pytest

from cases import rel_l2

pytestmark = pytest.mark.gpu


def _objects(config="circular_conebeam", shape=(16, 40, 32), u0=-30.0, v0=-24.0):
    from ct2 import phantoms, tomo3D
    sino = tomo3D.sinogram3D(config, {"radius": 500.0, "source_to_detector": 1000.0}, shape,
                             slen=2 * np.pi, ulen=shape[1] * 1.5, vlen=shape[2] * 1.5,
                             s0=0.0, u0=u0, v0=v0, s_include_endpoints=0)
    img = tomo3D.image3D((24, 24, 20), 24.0, 24.0, 20.0, -12.0, -12.0, -10.0)
    img.mat = phantoms.random_ellipsoids(24, 12.0, 8, 3)[:, :, 2:22].copy()
    return sino, img


@pytest.mark.parametrize("alg", ["linear", "nearest"])
def test_flip_detector_projection_and_backprojection(alg):
    sino, img = _objects()
    img.project_to(sino, alg="proj_" + alg)
    flip = sino.duplicate()
    flip.flip_detector()
    img.project_to(flip, alg="proj_" + alg)
    assert np.array_equal(flip.mat, sino.mat.transpose(0, 2, 1))
    from ct2 import tomo3D
    a = tomo3D.image3D((24, 24, 20), 24.0, 24.0, 20.0, -12.0, -12.0, -10.0)
    b = tomo3D.image3D((24, 24, 20), 24.0, 24.0, 20.0, -12.0, -12.0, -10.0)
    sino.backproject_to(a, alg="adj_" + alg)
    flip.backproject_to(b, alg="adj_" + alg)
    assert rel_l2(b.mat, a.mat) <= 1e-6


def test_flipped_geometry_equals_flipped_sinogram():
    sino, img = _objects()
    flip, _ = _objects("circular_conebeam_flipped", (16, 32, 40), u0=-24.0, v0=-30.0)
    img.project_to(sino)
    img.project_to(flip)
    assert np.array_equal(flip.mat, sino.mat.transpose(0, 2, 1))
