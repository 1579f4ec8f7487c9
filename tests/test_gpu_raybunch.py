"""Ray bunching on the GPU (ct2_sinoscale, ct2_sinoave, ct2_sinodist)
against the FP32 oracle, itself bit-identical to the compiled Fortran
(tomo3DfortranAux.f:1-176).  Bar: bit-identical."""
import numpy as np
import pytest

import oracle
from cases import cone_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def o32():
    return oracle.restated("f32")


@pytest.mark.parametrize("win", [(1, 2, 2, 1, 2, 2), (2, 3, 2, 3, 4, 2), (1, 4, 4, 1, 4, 4),
                                 (2, 2, 3, 1, 1, 1), (3, 5, 4, 3, 5, 4)])
def test_raybunch_bit_exact(o32, win):
    import torch
    from ct2 import tomo3Dfort
    ind, fv, geo, _, _ = cone_case(n=32, ns=30, nu=64, nv=60)
    ns, nu, nv, du, dv, u0, v0 = geo
    rng = np.random.default_rng(11)
    s = rng.standard_normal((ns, nu, nv)).astype(np.float32)
    a = tomo3Dfort.sinoscale(s, fv, ns, nu, nv, du, dv, u0, v0)
    assert np.array_equal(a, o32.tomo3Dfort.sinoscale(s, fv, ns, nu, nv, du, dv, u0, v0))
    nds, ndu, ndv, nws, nwu, nwv = win
    nss, nus, nvs = (ns - nws) // nds + 1, (nu - nwu) // ndu + 1, (nv - nwv) // ndv + 1
    n9 = (nss, nus, nvs, nds, ndu, ndv, nws, nwu, nwv)
    av = tomo3Dfort.sinoave(s, ns, nu, nv, *n9)
    assert np.array_equal(av, o32.tomo3Dfort.sinoave(s, ns, nu, nv, *n9))
    d = tomo3Dfort.sinodist(av, ns, nu, nv, *n9)
    assert np.array_equal(d, o32.tomo3Dfort.sinodist(av, ns, nu, nv, *n9))
    # device tensors stay on the device; sinoscale updates a float32 tensor in place
    t = torch.from_numpy(s.copy()).cuda()
    r = tomo3Dfort.sinoscale(t, fv, ns, nu, nv, du, dv, u0, v0)
    assert r.data_ptr() == t.data_ptr() and np.array_equal(t.cpu().numpy(), a)
    with pytest.raises(Exception):
        tomo3Dfort.sinoave(s, ns, nu, nv, nss + 1, nus, nvs, nds, ndu, ndv, nws, nwu, nwv)


def test_raybunch_objects(o32):
    from ct2.tomo3D import raybunch, sinogram3D
    s = sinogram3D("circular_conebeam", {"source_to_detector": 40., "radius": 20.}, (16, 32, 24),
                   ulen=8.0, vlen=6.0, u0=-4.0, v0=-3.0)
    s.mat = np.random.default_rng(4).random((16, 32, 24)).astype(np.float32)
    rb = raybunch(s, 2, 4, 4, 2, 4, 4)
    rb.get_data(s)
    assert rb.mat.shape == (8, 8, 6)
    want = o32.tomo3Dfort.sinoave(s.mat, 16, 32, 24, 8, 8, 6, 2, 4, 4, 2, 4, 4)
    assert np.array_equal(rb.mat, want)
    fine = s.duplicate()
    rb.put_data(fine)
    assert np.array_equal(fine.mat, o32.tomo3Dfort.sinodist(want, 16, 32, 24, 8, 8, 6, 2, 4, 4,
                                                            2, 4, 4))
    h = s.duplicate()
    h.homoscale()
    assert np.array_equal(h.mat, o32.tomo3Dfort.sinoscale(s.mat, s.frame_vectors, 16, 32, 24,
                                                          s.du, s.dv, s.u0, s.v0))
