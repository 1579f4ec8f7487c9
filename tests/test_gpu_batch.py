"""Batched plans (ct2.batch) against the single-problem operator.

Problems of uneven size -- cubic and non-cubic volumes, different view counts
and detectors, one z-driven tomosynthesis problem, rays switched off by the
indicator -- go through one persistent launch per operation.  Per problem:
the forward equals ConeBeamOperator.forward bit for bit (same device code);
the adjoint equals ConeBeamOperator.adjoint to LDS summation order (1e-6);
SART over interleaved subsets equals ct2_sart_step with view0 = s,
vstride = nsub applied problem by problem (1e-5); replays of a captured SART
iteration equal eager iterations.
"""
import numpy as np
import pytest
import torch

from cases import rel_l2

pytestmark = pytest.mark.gpu


def make_problems(arith, seed=11):
    from ct2 import configs3D, phantoms
    from ct2.operator import ConeBeamOperator
    rng = np.random.default_rng(seed)
    specs = [((20, 20, 20), 7), ((33, 33, 33), 25), ((24, 30, 18), 12), ((40, 40, 40), 16),
             ((27, 27, 27), 9)]
    ops, vols, sinos = [], [], []
    for i, (shape, ns) in enumerate(specs):
        n = max(shape)
        nd = int(np.ceil(1.5 * n))
        fv = configs3D.frame_vectors("circular_conebeam", {"radius": 300.0,
                                                          "source_to_detector": 600.0},
                                     ns, 2 * np.pi, 0.3 * i, 0)
        ind = (rng.random((ns, nd, nd + 3)) > 0.03).astype(np.int8) if i == 2 else None
        hw = np.array(shape) / 2.0
        op = ConeBeamOperator(fv, ns, nd, nd + 3, 1.0, 1.0, -nd / 2, -(nd + 3) / 2, shape,
                              (0.9, 0.9, 0.9), tuple(-0.9 * hw), arith=arith, indicator=ind)
        cube = phantoms.random_ellipsoids(n, 0.45 * n, 6, seed + i)[:shape[0], :shape[1], :shape[2]]
        v = torch.from_numpy(np.ascontiguousarray(cube)).cuda()
        ops.append(op)
        vols.append(v)
        sinos.append(op.forward(v))
    # a tomosynthesis problem, z-driven
    fv = configs3D.frame_vectors("tomosynthesis", {"radius": 650.0, "center_to_detector": 20.0},
                                 11, np.deg2rad(50), np.deg2rad(65), 1)
    op = ConeBeamOperator(fv, 11, 60, 50, 1.0, 1.0, -30.0, -25.0, (48, 40, 10), (0.8, 0.8, 2.0),
                          (-19.2, -16.0, 0.0), axis="z", arith=arith)
    v = torch.from_numpy(rng.random((48, 40, 10)).astype(np.float32)).cuda()
    ops.append(op)
    vols.append(v)
    sinos.append(op.forward(v))
    return ops, vols, sinos


@pytest.mark.parametrize("arith", ["exact", "fast"])
def test_batch_project_and_backproject(arith):
    from ct2.batch import BatchPlan
    ops, vols, sinos = make_problems(arith)
    proj = [torch.empty_like(s) for s in sinos]
    adj = [torch.empty_like(v) for v in vols]
    plan = BatchPlan(ops, sinos, vols, nsub=4, proj=proj, adj=adj, tune=True)
    info = plan.info()
    assert info["tile"][0] in (2, 4, 8) and info["nprob"] == len(ops)
    plan.project()
    plan.backproject()
    for i, op in enumerate(ops):
        ref = op.forward(vols[i])
        assert torch.equal(proj[i], ref), f"problem {i}: forward differs"
        refb = op.adjoint(sinos[i])
        e = rel_l2(adj[i].cpu().numpy(), refb.cpu().numpy())
        assert e < 1e-6, (i, e)
    plan.close()


@pytest.mark.parametrize("arith", ["exact", "fast"])
def test_batch_sart_matches_single_problem(arith):
    from ct2.batch import BatchPlan
    ops, vols, sinos = make_problems(arith, seed=12)
    nsub, lam, niter = 3, 0.8, 2
    xb = [torch.zeros_like(v) for v in vols]
    plan = BatchPlan(ops, sinos, xb, nsub=nsub, tune=False)
    plan.sart(lam, niter)
    for i, op in enumerate(ops):
        x = torch.zeros_like(vols[i])
        for _ in range(niter):
            for s in range(nsub):
                if s < op.sino_shape[0]:
                    op.sart_step(sinos[i], x, lam, view0=s, nviews=0, vstride=nsub)
        e = rel_l2(xb[i].cpu().numpy(), x.cpu().numpy())
        assert e < 1e-5, (i, e)
        # and SART approaches the truth
        assert rel_l2(x.cpu().numpy(), vols[i].cpu().numpy()) < 0.9
    plan.close()


def test_batch_sart_graph_replays():
    from ct2.batch import BatchPlan
    from ct2.solvers import capture
    ops, vols, sinos = make_problems("fast", seed=13)
    xe = [torch.zeros_like(v) for v in vols]
    xg = [torch.zeros_like(v) for v in vols]
    pe = BatchPlan(ops, sinos, xe, nsub=4, tune=False)
    pg = BatchPlan(ops, sinos, xg, nsub=4, tune=False)
    pg.sart(1.0, 1)
    g = capture(lambda: pg.sart(1.0, 1))
    pe.sart(1.0, 1)
    for k in range(3):
        g.replay()
        pe.sart(1.0, 1)
        torch.cuda.synchronize()
        for i in range(len(ops)):
            e = rel_l2(xg[i].cpu().numpy(), xe[i].cpu().numpy())
            assert e < 1e-5, (k, i, e)
    pe.close()
    pg.close()


def test_batch_rejects_bad_input():
    from ct2._lib import CT2Error
    from ct2.batch import BatchPlan
    ops, vols, sinos = make_problems("fast", seed=14)
    with pytest.raises(ValueError):
        BatchPlan(ops, sinos[:-1], vols)
    plan = BatchPlan(ops, sinos, vols, nsub=2, tune=False)
    with pytest.raises(CT2Error, match="no projection outputs"):
        plan.project()
    plan.close()
