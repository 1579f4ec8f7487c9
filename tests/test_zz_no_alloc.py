"""ct2_fdk_sino_filter3d_noalloc (the allocation-free variant of
ct2_fdk_sino_filter3d, no counterpart in the reference) through the C ABI:
two calls in a row, the second on a smaller sinogram with a shorter pad and
the other filter axis, each equal bit for bit to ct2_fdk_sino_filter3d on the
same input (same taps, same kernels) and within 1e-5 of the FDK oracle
(oracle/fdk.py, FDK.py:25-103); lines that pad beyond the static scratch are
rejected on the host."""
import numpy as np
import pytest
import torch

from oracle import fdk as ofdk
from cases import rel_l2

pytestmark = pytest.mark.gpu


def _filter(name, dp, dim, sino, hann=0, L=1.0):
    from ct2 import _dev
    from ct2._lib import call
    x = torch.from_numpy(sino).cuda()
    y = torch.full_like(x, float("nan"))
    ns, nu, nv = sino.shape
    call(name, float(dp), dim, ns, nu, nv, _dev.ptr(x), _dev.ptr(y), hann, float(L),
         _dev.stream_handle())
    torch.cuda.synchronize()
    return y.cpu().numpy()


@pytest.mark.parametrize("hann,L", [(0, 1.0), (1, 0.75)])
def test_noalloc_filter_matches_allocating_entry(hann, L):
    rng = np.random.default_rng(4)
    cases = [(0.8, 1, rng.normal(size=(6, 300, 40)).astype(np.float32)),   # npad 512
             (1.3, 2, rng.normal(size=(3, 20, 45)).astype(np.float32))]    # npad 128, along v
    for dp, dim, sino in cases:
        a = _filter("ct2_fdk_sino_filter3d_noalloc", dp, dim, sino, hann, L)
        b = _filter("ct2_fdk_sino_filter3d", dp, dim, sino, hann, L)
        assert np.array_equal(a, b)
        want = ofdk.fdk_sino_filter3D(dp, dim, sino.astype(np.float64),
                                      "Hann" if hann else "", L)
        assert rel_l2(a, want) <= 1e-5


def test_noalloc_filter_rejects_long_lines():
    from ct2._lib import CT2Error
    sino = np.zeros((1, 6000, 2), np.float32)   # 6000 bins pad to 16384 > 8192
    with pytest.raises(CT2Error, match="beyond"):
        _filter("ct2_fdk_sino_filter3d_noalloc", 1.0, 1, sino)
