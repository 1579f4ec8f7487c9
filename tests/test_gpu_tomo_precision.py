"""Precision study on the limited-angle tomosynthesis workload.

1024 x 768 x 64 volume (0.1 x 0.1 x 1 mm) of a breast-like phantom (200
background ellipsoids of density 0.2-0.35 plus 50 microcalcification spheres
of radius 0.1-0.3 mm at density 2.0, seed 20260103); 25 views over +-25 deg
(tomosynthesis geometry, R = 650 mm, detector 20 mm below the centre of
rotation); 2048 x 1536 detector at 0.1 mm; z-driven projector (the
reference's sinoprojz, view_proj3D_tomo.f:4-148) and its transpose.

Each output -- the forward projection with FP32, FP16 and BF16 volume
storage, and the adjoint written as FP32, FP16 and BF16 -- is scored by its
relative L2 error against the FP64 oracle on three views (first, middle,
last), and the reference-arithmetic outputs also against the Fortran-faithful
FP32 oracle.  The table is printed.  Bars: reference arithmetic reproduces
the reference (forward bit-exact, adjoint within 1e-6 of the FP32 oracle);
FP32 and FP16-storage forwards within 1e-5 and 1e-4 of FP64.  The adjoint's
distance to FP64 is dominated by the FP32 interpolation weights at voxel
indices up to 1024 (ulp 6e-5 there) with only three views contributing to a
voxel: about 1e-4 in the reference's own arithmetic, about 3e-5 with the
FP64 ray set-up of the fast arithmetic.
"""
import numpy as np
import pytest
import torch

import oracle
from cases import rel_l2

pytestmark = pytest.mark.gpu

NX, NY, NZ = 1024, 768, 64
VOX = (0.1, 0.1, 1.0)
NS, NU, NV, PITCH = 25, 2048, 1536, 0.1
VIEWS = [0, 12, 24]


@pytest.fixture(scope="module")
def problem():
    from ct2 import configs3D, phantoms
    fv = configs3D.frame_vectors("tomosynthesis", {"radius": 650.0, "center_to_detector": 20.0},
                                 NS, np.deg2rad(50.0), np.deg2rad(65.0), 1)
    vol = phantoms.breast_like((NX, NY, NZ), VOX, 20260103)
    geo = (NS, NU, NV, PITCH, PITCH, -NU * PITCH / 2, -NV * PITCH / 2)
    grid = (VOX[0], VOX[1], VOX[2], -NX * VOX[0] / 2, -NY * VOX[1] / 2, 0.0, NX, NY, NZ)
    return fv, vol, geo, grid


def test_tomosynthesis_precision_per_output(problem):
    from ct2.operator import ConeBeamOperator, to_storage
    fv, vol, geo, grid = problem
    ns, nu, nv, du, dv, u0, v0 = geo
    dx, dy, dz, x0, y0, z0, nx, ny, nz = grid
    ops = {a: ConeBeamOperator(fv, ns, nu, nv, du, dv, u0, v0, (nx, ny, nz), (dx, dy, dz),
                               (x0, y0, z0), axis="z", arith=a) for a in ("exact", "fast")}
    v32 = torch.from_numpy(vol).cuda()
    got = {}
    got["fwd f32 storage, reference arithmetic"] = ops["exact"].forward(v32)[VIEWS].cpu()
    got["fwd f32 storage, fast arithmetic"] = ops["fast"].forward(v32)[VIEWS].cpu()
    for name, dt in (("f16", torch.float16), ("bf16", torch.bfloat16)):
        for rnd in ("stochastic", "nearest"):
            v16 = to_storage(v32, dt, rnd)
            got[f"fwd {name} storage ({rnd} rounding), fast arithmetic"] = \
                ops["fast"].forward(v16)[VIEWS].cpu()
            del v16
    # adjoint of the FP32 forward on the checked views (zero elsewhere)
    y = torch.zeros(ops["fast"].sino_shape, dtype=torch.float32, device="cuda")
    y[VIEWS] = got["fwd f32 storage, reference arithmetic"].cuda()
    got["bp f32 out, reference arithmetic"] = ops["exact"].adjoint(y).cpu()
    got["bp f32 out, fast arithmetic"] = ops["fast"].adjoint(y).cpu()
    got["bp f16 out, fast arithmetic"] = ops["fast"].adjoint(y, dtype=torch.float16).cpu()
    got["bp bf16 out, fast arithmetic"] = ops["fast"].adjoint(y, dtype=torch.bfloat16).cpu()
    del y

    o64 = oracle.restated_raw("f64")
    o32 = oracle.restated_raw("f32")
    fv3 = np.ascontiguousarray(fv[VIEWS])
    geo3 = (len(VIEWS),) + tuple(geo[1:])
    ref_fwd = o64.sinoproj(oracle.LINEAR, 1, None, fv3, *geo3, vol, *grid)
    y3 = got["fwd f32 storage, reference arithmetic"].numpy()
    ref_bp = o64.sinobackproject(oracle.LINEAR, 1, y3, None, fv3, *geo3, *grid)
    f32_fwd = o32.sinoproj(oracle.LINEAR, 1, None, fv3, *geo3, vol, *grid)
    f32_bp = o32.sinobackproject(oracle.LINEAR, 1, y3, None, fv3, *geo3, *grid)
    assert np.array_equal(y3, f32_fwd)
    e32_bp = rel_l2(got["bp f32 out, reference arithmetic"].numpy(), f32_bp)
    print(f"\n  bp reference arithmetic vs FP32 oracle: {e32_bp:.3e}")
    assert e32_bp <= 1e-6

    err = {}
    for k, v in got.items():
        ref = ref_fwd if k.startswith("fwd") else ref_bp
        err[k] = rel_l2(v.float().numpy(), ref)
    print("\ntomosynthesis 1024x768x64, 25 views, 2048x1536: rel L2 vs FP64 oracle "
          f"on views {VIEWS}")
    for k, e in err.items():
        print(f"  {k:55s} {e:.3e}")
    assert err["fwd f32 storage, reference arithmetic"] <= 1e-5
    assert err["fwd f32 storage, fast arithmetic"] <= 1e-5
    assert err["fwd f16 storage (stochastic rounding), fast arithmetic"] <= 1e-4
    assert err["bp f32 out, reference arithmetic"] <= 2e-4
    assert err["bp f32 out, fast arithmetic"] <= 5e-5
    # 16-bit outputs are bounded by their own rounding (2^-11, 2^-8 relative)
    assert err["bp f16 out, fast arithmetic"] <= 5e-4
    assert err["bp bf16 out, fast arithmetic"] <= 4e-3
    assert err["fwd bf16 storage (stochastic rounding), fast arithmetic"] <= 4e-3
