"""Plane-major volume storage (CT2_*_ZPLANES, ct2_convert_volume_zplanes):
the volume as nz planes of C-ordered (x, y), for z-driven projection.

* The conversion: the float32 copy is the volume transposed to (nz, nx, ny)
  exactly, and the 16-bit copies equal the C-ordered conversion of the same
  rounding and seed, transposed, bit for bit (each element is rounded with
  its C-order index).
* The forward projection reads it through the grid's strides with the same
  samples and arithmetic, so it equals the projection of the C-ordered
  volume bit for bit: z-driven tomosynthesis (the reference's sinoprojz,
  view_proj3D_tomo.f:4-148) and circular cone-beam views with free axes, in
  fast and in reference arithmetic, f32/f16/bf16 storage, and the full
  1024 x 768 x 64 tomosynthesis volume on two views.
* Errors: a storage type that is not plane-major for the conversion, a bad
  shape, and the adjoint (which writes C-ordered volumes only).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _tomo_op(shape, ns=7, nu=128, nv=112, vox=(0.5, 0.5, 1.0), arith="fast", pitch=0.5):
    from ct2 import configs3D
    from ct2.operator import ConeBeamOperator
    nx, ny, nz = shape
    fv = configs3D.frame_vectors("tomosynthesis", {"radius": 650.0, "center_to_detector": 20.0},
                                 ns, np.deg2rad(50.0), np.deg2rad(65.0), 1)
    return ConeBeamOperator(fv, ns, nu, nv, pitch, pitch, -nu * pitch / 2, -nv * pitch / 2,
                            shape, vox, (-nx * vox[0] / 2, -ny * vox[1] / 2, 0.0), axis="z",
                            arith=arith)


def _cone_op(shape, ns=9, nd=60, arith="fast"):
    from ct2 import configs3D
    from ct2.operator import ConeBeamOperator
    fv = configs3D.frame_vectors("circular_conebeam", {"radius": 300.0, "source_to_detector": 600.0},
                                 ns, 2 * np.pi, 0.3, 0)
    hw = np.array(shape) / 2.0
    return ConeBeamOperator(fv, ns, nd, nd + 4, 1.5, 1.5, -nd * 0.75, -(nd + 4) * 0.75, shape,
                            (1.0, 1.0, 1.0), tuple(-hw), arith=arith)


def _volume(shape, seed):
    from ct2 import phantoms
    n = max(shape)
    v = phantoms.random_ellipsoids(n, 0.45 * n, 12, seed)[:shape[0], :shape[1], :shape[2]]
    return torch.from_numpy(np.ascontiguousarray(v, dtype=np.float32)).cuda()


def test_conversion_is_the_transposed_c_order_conversion():
    from ct2.operator import to_storage
    v = _volume((37, 29, 23), 1)
    z32 = to_storage(v, torch.float32, zplanes=True)
    assert tuple(z32.shape) == (23, 37, 29)
    assert torch.equal(z32, v.permute(2, 0, 1).contiguous())
    for dt in (torch.float16, torch.bfloat16):
        for rnd in ("stochastic", "nearest"):
            c = to_storage(v, dt, rnd, seed=99)
            z = to_storage(v, dt, rnd, seed=99, zplanes=True)
            assert torch.equal(z.view(torch.int16), c.permute(2, 0, 1).contiguous().view(torch.int16)), (dt, rnd)


@pytest.mark.parametrize("arith", ["fast", "exact"])
@pytest.mark.parametrize("geometry", ["tomosynthesis", "cone"])
def test_forward_equals_c_order_forward(arith, geometry):
    from ct2.operator import to_storage
    if geometry == "tomosynthesis":
        shape = (96, 80, 16)
        op = _tomo_op(shape, arith=arith)
    else:
        shape = (30, 34, 26)
        op = _cone_op(shape, arith=arith)
    v = _volume(shape, 2)
    for dt in (torch.float32, torch.float16, torch.bfloat16):
        c = v if dt == torch.float32 else to_storage(v, dt, "nearest")
        z = to_storage(v, dt, "nearest", zplanes=True)
        a = op.forward(c)
        b = op.forward(z, layout="zxy")
        assert torch.isfinite(b).all() and float(b.abs().max()) > 0
        assert torch.equal(a, b), (dt, float((a - b).abs().max()))


def test_full_size_tomosynthesis_two_views():
    """1024 x 768 x 64 (0.1 x 0.1 x 1 mm), 2048 x 1536 detector at 0.1 mm:
    views 0 and 24 of the 25, f16 storage, fast arithmetic."""
    from ct2.operator import to_storage
    shape = (1024, 768, 64)
    op = _tomo_op(shape, ns=25, nu=2048, nv=1536, vox=(0.1, 0.1, 1.0), pitch=0.1)
    g = torch.Generator(device="cuda").manual_seed(5)
    v = torch.rand(shape, device="cuda", generator=g)
    c = to_storage(v, torch.float16, "stochastic")
    z = to_storage(v, torch.float16, "stochastic", zplanes=True)
    del v
    a = op.forward(c, view0=0, nviews=2, vstride=24)
    b = op.forward(z, view0=0, nviews=2, vstride=24, layout="zxy")
    assert torch.equal(a[[0, 24]], b[[0, 24]])
    assert float(b[[0, 24]].abs().max()) > 0


def test_rejections():
    from ct2 import _dev
    from ct2._lib import CT2_F16, CT2_F32_ZPLANES, CT2Error, call
    v = _volume((8, 8, 8), 3)
    out = torch.empty(512, dtype=torch.float16, device="cuda")
    with pytest.raises(CT2Error, match="destination must be"):
        call("ct2_convert_volume_zplanes", _dev.ptr(v), _dev.ptr(out), CT2_F16, 8, 8, 8, 0, 0,
             _dev.stream_handle())
    with pytest.raises(CT2Error, match="bad shape"):
        call("ct2_convert_volume_zplanes", _dev.ptr(v), _dev.ptr(out), CT2_F32_ZPLANES, 0, 8, 8,
             0, 0, _dev.stream_handle())
    # the adjoint writes C-ordered volumes only
    op = _cone_op((8, 8, 8), nd=12)
    d = op._desc(CT2_F32_ZPLANES)
    y = torch.ones(op.sino_shape, device="cuda")
    with pytest.raises(CT2Error, match="vol_dtype"):
        call("ct2_backproject", C.byref(d), _dev.ptr(v), _dev.ptr(y), None, _dev.ptr(op.fv),
             _dev.stream_handle())
