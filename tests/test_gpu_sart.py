"""SART on the GPU against the oracle composition (oracle/sart.py), and the
batched SART at the full 64-problem size.

* ct2_sart_step (ConeBeamOperator.sart_step) over view subsets, in exact and
  fast arithmetic, linear and nearest sampling, with indicator holes, rays
  that miss the volume (zero row sums) and voxels no used ray reaches (zero
  column sums): relative L2 <= 1e-5 against the oracle composition, in FP32
  for exact arithmetic (the reference's FP32 operation order) and in FP64
  for fast arithmetic.
* A one-problem batched plan (ct2_batch_sart): the same bar over nsub
  interleaved subsets and several iterations.
* 64 problems drawn as the batching workload draws them (seed 20260104:
  cubic n ~ U{64..192}, views ~ U{30..180} over 360 deg, detector
  ceil(1.5 n)^2, random-ellipsoid phantoms), 10 SART iterations with the
  autotuned plan: three sampled problems equal ct2_sart_step applied problem
  by problem bit for bit (the plan's adjoint runs every problem on the
  separable tiles), every problem's data residual falls each iteration, and
  a HIP graph of the 10 iterations replays to the eager result.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from oracle import sart as osart
from cases import cone_case, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def o64():
    return oracle.restated_raw("f64")


def _oracle_for(arith):
    """Exact arithmetic is the reference's FP32 operation order: its oracle is
    the FP32 restatement (same forward bit for bit, same float32 residual
    arithmetic); fast arithmetic is held against the FP64 restatement."""
    return oracle.restated_raw("f32" if arith == "exact" else "f64")


def _op(ind, fv, geo, g2, arith, interp="linear"):
    from ct2.operator import ConeBeamOperator
    ns, nu, nv, du, dv, u0, v0 = geo
    dx, dy, dz, x0, y0, z0, nx, ny, nz = g2
    return ConeBeamOperator(fv, ns, nu, nv, du, dv, u0, v0, (nx, ny, nz), (dx, dy, dz),
                            (x0, y0, z0), interp=interp, arith=arith, indicator=ind)


def _guard_case():
    """Detector shifted in u and tall in v: some rays miss the volume, and
    the first view's subset leaves voxels unseen."""
    ind, fv, geo, vol, g2 = cone_case(n=24, ns=10, nu=20, nv=40, seed=3)
    ns, nu, nv, du, dv, u0, v0 = geo
    return ind, fv, (ns, nu, nv, du, dv, u0 + 9.0, v0), vol, g2


@pytest.mark.parametrize("arith", ["exact", "fast"])
@pytest.mark.parametrize("interp", ["linear", "nearest"])
@pytest.mark.parametrize("view0,nviews,vstride", [(0, 0, 1), (1, 0, 3), (0, 1, 1), (4, 3, 2)])
def test_sart_step_vs_oracle(o64, arith, interp, view0, nviews, vstride):
    ind, fv, geo, vol, g2 = _guard_case()
    rng = np.random.default_rng(7)
    b = o64.sinoproj(0 if interp == "linear" else 1, 0, ind, fv, *geo, vol, *g2)
    b = (b * (1 + 0.02 * rng.standard_normal(b.shape))).astype(np.float32)
    x0 = (0.5 * vol + 0.1 * rng.random(vol.shape)).astype(np.float32)
    op = _op(ind, fv, geo, g2, arith, interp)
    x = torch.from_numpy(x0).cuda()
    op.sart_step(torch.from_numpy(b).cuda(), x, 0.8, view0=view0, nviews=nviews,
                 vstride=vstride)
    want = osart.sart_step(_oracle_for(arith), 0 if interp == "linear" else 1, 0, 0.8, b, ind,
                           fv, *geo, x0, *g2, view0=view0, nviews=nviews, vstride=vstride)
    got = x.cpu().numpy()
    assert np.isfinite(got).all()
    # the change, not only the result, must match (x0 dominates the norm)
    e = rel_l2(got - x0, want - x0)
    assert e <= 1e-5, e
    assert rel_l2(got, want) <= 1e-5


@pytest.mark.parametrize("arith", ["exact", "fast"])
def test_batch_sart_one_problem_vs_oracle(o64, arith):
    from ct2.batch import BatchPlan
    ind, fv, geo, vol, g2 = _guard_case()
    b = o64.sinoproj(0, 0, ind, fv, *geo, vol, *g2).astype(np.float32)
    op = _op(ind, fv, geo, g2, arith)
    x = torch.zeros(vol.shape, dtype=torch.float32, device="cuda")
    plan = BatchPlan([op], [torch.from_numpy(b).cuda()], [x], nsub=3, tune=False)
    plan.sart(0.9, 3)
    want = osart.sart_sweeps(_oracle_for(arith), 0, 0, 0.9, 3, 3, b, ind, fv, *geo,
                             np.zeros(vol.shape), *g2)
    e = rel_l2(x.cpu().numpy(), want)
    plan.close()
    assert e <= 1e-5, e


def _c4_problems():
    from ct2 import configs3D, phantoms
    from ct2.operator import ConeBeamOperator
    rng = np.random.default_rng(20260104)
    ops, truth, sinos = [], [], []
    for p in range(64):
        n = int(rng.integers(64, 193))
        ns = int(rng.integers(30, 181))
        nd = int(np.ceil(1.5 * n))
        fv = configs3D.frame_vectors("circular_conebeam", {"radius": 500.0,
                                                          "source_to_detector": 1000.0},
                                     ns, 2 * np.pi, 0.0, 0)
        op = ConeBeamOperator(fv, ns, nd, nd, 1.6, 1.6, -nd * 0.8, -nd * 0.8, (n, n, n),
                              (1.0, 1.0, 1.0), (-n / 2, -n / 2, -n / 2), arith="fast")
        v = phantoms.random_ellipsoids_device((n, n, n), 1.0, 20, rng)
        ops.append(op)
        truth.append(v)
        sinos.append(op.forward(v))
    return ops, truth, sinos


def test_c4_batched_sart_full_size():
    from ct2.batch import BatchPlan
    from ct2.solvers import capture
    ops, truth, sinos = _c4_problems()
    sizes = [op.shape[0] for op in ops]
    assert min(sizes) >= 64 and max(sizes) <= 192 and len(ops) == 64
    nsub, niter, lam = 10, 10, 1.0
    x = [torch.zeros_like(v) for v in truth]
    proj = [torch.empty_like(s) for s in sinos]
    plan = BatchPlan(ops, sinos, x, nsub=nsub, proj=proj, tune=True)
    info = plan.info()
    assert info["tile"][0] in (2, 4, 8) and info["nprob"] == 64
    res_prev = [float(torch.linalg.vector_norm(s)) for s in sinos]
    for it in range(niter):
        plan.sart(lam, 1)
        plan.project()
        res = [float(torch.linalg.vector_norm(s - q)) for s, q in zip(sinos, proj)]
        falling = [r < rp for r, rp in zip(res, res_prev)]
        assert all(falling), (it, [i for i, f in enumerate(falling) if not f])
        res_prev = res
    print(f"C4: mean data residual after {niter} iterations "
          f"{np.mean([r / float(torch.linalg.vector_norm(s)) for r, s in zip(res, sinos)]):.4f}",
          flush=True)
    # sampled problems against ct2_sart_step applied problem by problem: every
    # problem's views have their rows along z, so the plan's adjoint runs on
    # the separable tiles and each step is ct2_sart_step's, bit for bit
    order = np.argsort(sizes)
    for i in (int(order[0]), int(order[32]), int(order[-1])):
        xs = torch.zeros_like(truth[i])
        for _ in range(niter):
            for s in range(nsub):
                ops[i].sart_step(sinos[i], xs, lam, view0=s, nviews=0, vstride=nsub)
        e = rel_l2(x[i].cpu().numpy(), xs.cpu().numpy())
        assert torch.equal(x[i], xs), (i, sizes[i], e)
    # a HIP graph of the whole 10-iteration loop replays to the eager result
    xg = [torch.zeros_like(v) for v in truth]
    pg = BatchPlan(ops, sinos, xg, nsub=nsub, tune=False)
    pg.sart(lam, 1)
    g = capture(lambda: pg.sart(lam, niter))
    for v in xg:
        v.zero_()
    g.replay()
    torch.cuda.synchronize()
    worst = max(rel_l2(a.cpu().numpy(), b.cpu().numpy()) for a, b in zip(xg, x))
    assert worst < 1e-5, worst
    plan.close()
    pg.close()
