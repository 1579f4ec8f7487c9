"""HIP Joseph projector and its adjoint against the oracle.

Forward (exact arithmetic): bit-identical to the FP32 oracle, which is itself
bit-identical to the reference Fortran.  Adjoint: same matrix entries, only
the summation order differs -> relative L2 <= 1e-6 against the oracle, and
the dot-product test <Ax,y> = <x,A^T y>.
"""
import numpy as np
import pytest

import oracle
from cases import cone_case, ray_case, rel_l2, tomo_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def o32():
    return oracle.restated("f32")


@pytest.fixture(scope="module")
def o64():
    return oracle.restated("f64")


@pytest.mark.parametrize("interp", ["linear", "nearest"])
@pytest.mark.parametrize("case", [dict(), dict(n=24, ns=9, nu=33, nv=21, phantom="rand", seed=3),
                                  dict(n=20, ns=5, nu=64, nv=70, du=0.7, dv=0.6, seed=7)])
def test_forward_bit_exact(o32, interp, case):
    from ct2 import view_proj3D as vp
    ind, fv, geo, vol, g2 = cone_case(**case)
    a = getattr(vp, f"sinoproj_{interp}")(ind, fv, *geo, vol, *g2)
    b = getattr(o32.view_proj3D, f"sinoproj_{interp}")(ind, fv, *geo, vol, *g2)
    assert a.dtype == np.float32 and a.shape == b.shape
    assert np.abs(b).max() > 1.0
    assert np.array_equal(a, b), f"max diff {np.abs(a - b).max()}"


def test_forward_tomo_bit_exact(o32):
    from ct2 import view_proj3D_tomo as vpt
    ind, fv, geo, vol, g2 = tomo_case()
    a = vpt.sinoprojz(ind, fv, *geo, vol, *g2)
    b = o32.view_proj3D_tomo.sinoprojz(ind, fv, *geo, vol, *g2)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("interp", ["linear", "nearest"])
def test_forward_raylist_bit_exact(o32, interp):
    from ct2 import ray_proj3D as rp
    ind, rv, nr, vol, g3 = ray_case()
    a = getattr(rp, f"sinoproj_{interp}")(ind, rv, nr, vol, *g3)
    b = getattr(o32.ray_proj3D, f"sinoproj_{interp}")(ind, rv, nr, vol, *g3)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("interp", ["linear", "nearest"])
@pytest.mark.parametrize("case", [dict(), dict(n=24, ns=9, nu=33, nv=21, phantom="rand", seed=3),
                                  dict(n=40, ns=4, nu=100, nv=90, du=0.9, dv=0.9, seed=11)])
def test_adjoint_matches_oracle(o32, interp, case):
    from ct2 import view_proj3D as vp
    ind, fv, geo, vol, g2 = cone_case(**case)
    rng = np.random.default_rng(1)
    sino = rng.standard_normal((geo[0], geo[1], geo[2])).astype(np.float32)
    a = getattr(vp, f"sinobackproject_{interp}")(sino, ind, fv, *geo, *g2)
    b = getattr(o32.view_proj3D, f"sinobackproject_{interp}")(sino, ind, fv, *geo, *g2)
    assert np.abs(b).max() > 0.1
    assert rel_l2(a, b) <= 1e-6


def test_adjoint_raylist_matches_oracle(o32):
    from ct2 import ray_proj3D as rp
    ind, rv, nr, vol, g3 = ray_case()
    s = np.random.default_rng(2).standard_normal(nr).astype(np.float32)
    a = rp.sinobackproject_nearest(s, ind, rv, nr, *g3)
    b = o32.ray_proj3D.sinobackproject_nearest(s, ind, rv, nr, *g3)
    assert rel_l2(a, b) <= 1e-6


@pytest.mark.parametrize("interp", ["linear", "nearest"])
def test_dot_product(interp):
    """<Ax, y> == <x, A^T y> for random Gaussian x, y (sums in float64)."""
    from ct2 import view_proj3D as vp
    ind, fv, geo, vol, g2 = cone_case(n=32, ns=16, nu=48, nv=44)
    rng = np.random.default_rng(20260106)
    x = rng.standard_normal(vol.shape).astype(np.float32)
    y = rng.standard_normal((geo[0], geo[1], geo[2])).astype(np.float32)
    Ax = getattr(vp, f"sinoproj_{interp}")(ind, fv, *geo, x, *g2)
    Aty = getattr(vp, f"sinobackproject_{interp}")(y, ind, fv, *geo, *g2)
    lhs = float(np.dot(Ax.ravel().astype(np.float64), y.ravel().astype(np.float64)))
    rhs = float(np.dot(x.ravel().astype(np.float64), Aty.ravel().astype(np.float64)))
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), abs(rhs))


def test_adjoint_tomo_dot_product():
    """z-forced adjoint (no reference counterpart) is the transpose of sinoprojz."""
    import torch
    from ct2 import _dev
    from ct2._lib import call, ct2_desc, CT2_AXIS_Z
    import ctypes as C
    ind, fv, geo, vol, g2 = tomo_case()
    ns, nu, nv, du, dv, u0, v0 = geo
    dx, dy, dz, x0, y0, z0, nx, ny, nz = g2
    d = ct2_desc()
    call("ct2_desc_init", C.byref(d), ns, nu, nv, du, dv, u0, v0, nx, ny, nz, dx, dy, dz, x0, y0, z0)
    d.axis = CT2_AXIS_Z
    rng = np.random.default_rng(4)
    x = torch.from_numpy(rng.standard_normal((nx, ny, nz)).astype(np.float32)).cuda()
    y = torch.from_numpy(rng.standard_normal((ns, nu, nv)).astype(np.float32)).cuda()
    fvd = torch.from_numpy(fv).cuda()
    Ax = torch.empty_like(y)
    Aty = torch.empty_like(x)
    s = _dev.stream_handle()
    call("ct2_project", C.byref(d), _dev.ptr(Ax), None, _dev.ptr(fvd), _dev.ptr(x), s)
    call("ct2_backproject", C.byref(d), _dev.ptr(Aty), _dev.ptr(y), None, _dev.ptr(fvd), s)
    lhs = float((Ax.double() * y.double()).sum())
    rhs = float((x.double() * Aty.double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), abs(rhs))


@pytest.mark.parametrize("dtype,tol", [("f32", 2e-6), ("f16", 1e-3), ("bf16", 8e-3)])
def test_fast_and_mixed_forward_vs_fp64(o64, dtype, tol):
    """Fast arithmetic (FP64 setup, FMA stepping, Kahan sums) and 16-bit
    volume storage: relative L2 against the FP64 oracle within tol."""
    import ctypes as C
    import torch
    from ct2 import _dev
    from ct2._lib import call, ct2_desc, CT2_ARITH_FAST, CT2_F16, CT2_BF16, CT2_F32
    ind, fv, geo, vol, g2 = cone_case(n=48, ns=10, nu=72, nv=64, du=1.0, dv=1.0, holes=False)
    ns, nu, nv, du, dv, u0, v0 = geo
    dx, dy, dz, x0, y0, z0, nx, ny, nz = g2
    d = ct2_desc()
    call("ct2_desc_init", C.byref(d), ns, nu, nv, du, dv, u0, v0, nx, ny, nz, dx, dy, dz, x0, y0, z0)
    d.arith = CT2_ARITH_FAST
    tdt = {"f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16}[dtype]
    d.vol_dtype = {"f32": CT2_F32, "f16": CT2_F16, "bf16": CT2_BF16}[dtype]
    v = torch.from_numpy(vol).cuda().to(tdt)
    fvd = torch.from_numpy(fv).cuda()
    out = torch.empty((ns, nu, nv), dtype=torch.float32, device="cuda")
    call("ct2_project", C.byref(d), _dev.ptr(out), None, _dev.ptr(fvd), _dev.ptr(v), _dev.stream_handle())
    ref = o64.view_proj3D.sinoproj_linear(None, fv, *geo, vol, *g2)
    err = rel_l2(out.cpu().numpy(), ref)
    print(f"{dtype}: rel L2 vs FP64 = {err:.3e}")
    assert err <= tol


def _desc(geo, g2, **kw):
    import ctypes as C
    from ct2._lib import call, ct2_desc
    ns, nu, nv, du, dv, u0, v0 = geo
    dx, dy, dz, x0, y0, z0, nx, ny, nz = g2
    d = ct2_desc()
    call("ct2_desc_init", C.byref(d), ns, nu, nv, du, dv, u0, v0, nx, ny, nz, dx, dy, dz, x0, y0, z0)
    for k, v in kw.items():
        setattr(d, k, v)
    return d


@pytest.mark.parametrize("interp", [0, 1])
def test_fast_pair_dot_product(interp):
    """Fast-arithmetic forward and the brick adjoint share their matrix
    entries (a named brick keeps the adjoint on the brick kernel; the
    column-separable route has its own tests)."""
    import ctypes as C
    import torch
    from ct2 import _dev
    from ct2._lib import call, CT2_ARITH_FAST
    ind, fv, geo, vol, g2 = cone_case(n=40, ns=12, nu=64, nv=56, du=1.0, dv=1.0)
    d = _desc(geo, g2, arith=CT2_ARITH_FAST, interp=interp)
    d.brick[0], d.brick[1], d.brick[2] = 10, 10, 24
    rng = np.random.default_rng(31)
    x = torch.from_numpy(rng.standard_normal(vol.shape).astype(np.float32)).cuda()
    y = torch.from_numpy(rng.standard_normal(geo[:3]).astype(np.float32)).cuda()
    fvd = torch.from_numpy(fv).cuda()
    indd = torch.from_numpy(ind).cuda()
    Ax, Aty = torch.empty_like(y), torch.empty_like(x)
    s = _dev.stream_handle()
    call("ct2_project", C.byref(d), _dev.ptr(Ax), _dev.ptr(indd), _dev.ptr(fvd), _dev.ptr(x), s)
    call("ct2_backproject", C.byref(d), _dev.ptr(Aty), _dev.ptr(y), _dev.ptr(indd), _dev.ptr(fvd), s)
    lhs = float((Ax.double() * y.double()).sum())
    rhs = float((x.double() * Aty.double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), abs(rhs))


@pytest.mark.parametrize("brick", [(0, 0, 0), (8, 8, 8), (16, 8, 32), (5, 7, 9)])
def test_fast_adjoint_vs_fp64(o64, brick):
    """Fast adjoint against the FP64 oracle's exact adjoint, several brick shapes."""
    import ctypes as C
    import torch
    from ct2 import _dev
    from ct2._lib import call, CT2_ARITH_FAST
    ind, fv, geo, vol, g2 = cone_case(n=36, ns=10, nu=60, nv=50, du=1.0, dv=1.0)
    d = _desc(geo, g2, arith=CT2_ARITH_FAST)
    d.brick[0], d.brick[1], d.brick[2] = brick
    rng = np.random.default_rng(32)
    y = rng.standard_normal(geo[:3]).astype(np.float32)
    out = torch.empty(vol.shape, dtype=torch.float32, device="cuda")
    # keep every device tensor referenced until the kernel has run: a freed
    # temporary's block is reused by the caching allocator
    yd, indd, fvd = (torch.from_numpy(a).cuda() for a in (y, ind, fv))
    call("ct2_backproject", C.byref(d), _dev.ptr(out), _dev.ptr(yd), _dev.ptr(indd),
         _dev.ptr(fvd), _dev.stream_handle())
    torch.cuda.synchronize()
    ref = o64.view_proj3D.sinobackproject_linear(y, ind, fv, *geo, *g2)
    err = rel_l2(out.cpu().numpy(), ref)
    assert err <= 2e-5, err


@pytest.mark.parametrize("rounding", ["stochastic", "nearest"])
@pytest.mark.parametrize("interp,case", [("linear", "cone"), ("nearest", "cone"),
                                         ("linear", "tomo")])
def test_f16_pair_storage_projects_as_f16(rounding, interp, case):
    """The f16 pair storage (each value with its successor along z) holds the
    float16 conversion's values, and the fast forward projects it bit for bit
    as it projects the float16 volume: along z (cone beam, x- or y-driven
    rays) with the paired loads, across z (z-driven tomosynthesis) with the
    values alone."""
    import torch
    from ct2.operator import ConeBeamOperator, to_storage
    if case == "cone":
        ind, fv, geo, vol, g2 = cone_case(n=40, ns=8, nu=60, nv=52, du=1.0, dv=1.0)
        axis = "free"
    else:
        ind, fv, geo, vol, g2 = tomo_case()
        axis = "z"
    ns, nu, nv, du, dv, u0, v0 = geo
    dx, dy, dz, x0, y0, z0, nx, ny, nz = g2
    op = ConeBeamOperator(fv, ns, nu, nv, du, dv, u0, v0, (nx, ny, nz), (dx, dy, dz),
                          (x0, y0, z0), interp=interp, axis=axis, arith="fast", indicator=ind)
    v32 = torch.from_numpy(vol).cuda()
    h = to_storage(v32, torch.float16, rounding, seed=7)
    hp = to_storage(v32, torch.float16, rounding, seed=7, pairs=True)
    assert hp.shape == tuple(vol.shape) + (2,)
    flat, pf = h.reshape(-1), hp.reshape(-1, 2)
    assert torch.equal(pf[:, 0].view(torch.int16), flat.view(torch.int16))
    assert torch.equal(pf[:-1, 1].view(torch.int16), flat[1:].view(torch.int16))
    assert int(pf[-1, 1].view(torch.int16)) == 0
    a = op.forward(h).cpu().numpy()
    b = op.forward(hp).cpu().numpy()
    assert np.abs(a).max() > 0.5
    assert np.array_equal(a, b), np.abs(a - b).max()
