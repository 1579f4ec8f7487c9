"""ct2.configs3D against the reference's own geometry functions: frame vectors
of circular_conebeam, circular_conebeam_flipped and tomosynthesis
(emil_code/configs3D.py:11-49, :196-234, :615-654) evaluated at the views of
sinogram3D.__init__ with get_ds (tomo3D.py:113-147, :186-190), and the
ball_flatpanel ray vectors (linoConfigs3D.py:89-135), as written by
tests/golden/make_geometry.py from the reference modules themselves.  Bit for
bit after the float32 cast the kernels read."""
import os

import numpy as np
import pytest

from ct2 import configs3D

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "geometry_reference.npz")


def _cases():
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "make_geometry", os.path.join(HERE, "golden", "make_geometry.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.CASES, mod.LINO_PARMS


CASES, LINO_PARMS = _cases()


@pytest.mark.parametrize("q", range(len(CASES)))
def test_frame_vectors_bit_exact(q):
    name, parms, ns, slen, s0, ends = CASES[q]
    want = np.load(GOLD)[f"case{q}_fv"]
    got = configs3D.frame_vectors(name, parms, ns, slen, s0, ends)
    assert got.dtype == np.float32 and got.shape == (ns, 18)
    assert np.array_equal(got, want), np.abs(got - want).max()


def test_lino_ray_vectors_bit_exact_as_float32():
    g = np.load(GOLD)
    got = np.array([configs3D.lino_spherical_flatpanel(tuple(s), LINO_PARMS) for s in g["lino_s"]])
    assert np.array_equal(got.astype(np.float32), g["lino_rv"].astype(np.float32))
    assert np.allclose(got, g["lino_rv"], rtol=1e-15, atol=1e-12)
