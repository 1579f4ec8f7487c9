"""Batched plans (ct2.batch) against the single-problem operator.

Problems of uneven size -- cubic and non-cubic volumes, different view counts
and detectors, one z-driven tomosynthesis problem, rays switched off by the
indicator -- go through one persistent launch per operation.  Per problem:
the forward equals ConeBeamOperator.forward bit for bit (same device code);
the adjoint equals ConeBeamOperator.adjoint to LDS summation order (1e-6);
SART over interleaved subsets equals ct2_sart_step with view0 = s,
vstride = nsub applied problem by problem (1e-5); replays of a captured SART
iteration equal eager iterations.  A plan of one problem on the
column-separable kernels gives ct2_sart_step's updates bit for bit, and agrees
with the batched kernels (CT2_PLAN_SEP=0) to FP32 summation order; in plans of
several problems, those whose SART adjoint runs on the batched separable tiles
(every z-cell class) give ct2_sart_step's updates bit for bit as well.
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


# ---- one-problem plans on the column-separable kernels ----
def _sep_problem(shape=(36, 28, 40), ns=20, nu=50, nv=44, tilt_every=0, seed=21):
    from ct2 import configs3D, phantoms
    from ct2.operator import ConeBeamOperator
    rng = np.random.default_rng(seed)
    fv = configs3D.frame_vectors("circular_conebeam", {"radius": 300.0, "source_to_detector": 600.0},
                                 ns, 2 * np.pi, 0.2, 0).astype(np.float64)
    for p in range(0, ns, tilt_every or ns + 1):
        # e_v leans 2 degrees towards e_u: that view's rows no longer run along z
        eu, ev = fv[p, 9:12].copy(), fv[p, 12:15].copy()
        a = np.deg2rad(2.0)
        fv[p, 12:15] = np.cos(a) * ev + np.sin(a) * eu
        fv[p, 9:12] = np.cos(a) * eu - np.sin(a) * ev
    ind = (rng.random((ns, nu, nv)) > 0.04).astype(np.int8)
    hw = np.array(shape) / 2.0
    op = ConeBeamOperator(fv.astype(np.float32), ns, nu, nv, 1.6, 1.6, -nu * 0.8, -nv * 0.8,
                          shape, (1.0, 1.0, 1.0), tuple(-hw), arith="fast", indicator=ind)
    n = max(shape)
    cube = phantoms.random_ellipsoids(n, 0.45 * n, 6, seed)[:shape[0], :shape[1], :shape[2]]
    v = torch.from_numpy(np.ascontiguousarray(cube)).cuda()
    return op, v, op.forward(v)


def _sart_steps(op, sino, nsub, lam, niter):
    x = torch.zeros(op.shape, dtype=torch.float32, device="cuda")
    for _ in range(niter):
        for s in range(min(nsub, op.sino_shape[0])):
            op.sart_step(sino, x, lam, view0=s, nviews=0, vstride=nsub)
    return x


@pytest.mark.parametrize("case", ["plain", "tilted_views", "more_subsets_than_views"])
def test_one_problem_plan_sart_equals_sart_step(case):
    """A one-problem plan on the separable kernels gives ct2_sart_step's
    subset updates bit for bit (the same forward, residual, adjoint, row and
    column sums); tilted views go through the brick pass of the adjoint in
    both."""
    from ct2.batch import BatchPlan
    kw = {"plain": {}, "tilted_views": dict(tilt_every=3),
          "more_subsets_than_views": dict(ns=3)}[case]
    op, v, sino = _sep_problem(**kw)
    nsub, lam, niter = 4, 0.9, 2
    x = torch.zeros_like(v)
    plan = BatchPlan([op], [sino], [x], nsub=nsub)
    info = plan.info()
    nvox, nsino = v.numel(), sino.numel()
    # the route's extra memory: a correction volume and one subset's tables
    assert info["bytes"] > 4 * (2 * nsino + nvox * nsub) + 4 * nvox
    plan.sart(lam, niter)
    ref = _sart_steps(op, sino, nsub, lam, niter)
    assert torch.equal(x, ref)
    assert rel_l2(x.cpu().numpy(), v.cpu().numpy()) < 0.9
    plan.close()


def test_one_problem_plan_routes_agree(monkeypatch):
    """CT2_PLAN_SEP=0 keeps the batched kernels for the one-problem plan:
    the two routes agree to FP32 summation order."""
    from ct2.batch import BatchPlan
    op, v, sino = _sep_problem(seed=22)
    xs, xb = torch.zeros_like(v), torch.zeros_like(v)
    ps = BatchPlan([op], [sino], [xs], nsub=5, tune=False)
    monkeypatch.setenv("CT2_PLAN_SEP", "0")
    pb = BatchPlan([op], [sino], [xb], nsub=5, tune=False)
    ps.sart(1.0, 2)
    pb.sart(1.0, 2)
    e = rel_l2(xs.cpu().numpy(), xb.cpu().numpy())
    assert 0 < e < 1e-5, e
    assert pb.info()["bytes"] < ps.info()["bytes"]
    ps.close()
    pb.close()


def test_one_problem_plan_graph_replays_with_device_lambda():
    """ct2_batch_sart_dev on the separable route, captured once and replayed
    with the relaxation changed on the device between replays, equals the
    eager calls bit for bit."""
    from ct2 import _dev
    from ct2._lib import call
    from ct2.batch import BatchPlan
    from ct2.solvers import capture
    op, v, sino = _sep_problem(shape=(32, 32, 48), ns=30, nu=56, nv=60, seed=23)
    xe, xg = torch.zeros_like(v), torch.zeros_like(v)
    pe = BatchPlan([op], [sino], [xe], nsub=6, tune=False)
    pg = BatchPlan([op], [sino], [xg], nsub=6, tune=False)
    lam = torch.ones((1,), dtype=torch.float32, device="cuda")

    def step(plan):
        call("ct2_batch_sart_dev", plan._h, _dev.ptr(lam), 1, _dev.stream_handle())

    lam.fill_(1.0)
    step(pg)
    g = capture(lambda: step(pg))
    step(pe)
    for k, val in enumerate((0.8, 0.6, 0.5)):
        lam.fill_(val)
        g.replay()
        step(pe)
        torch.cuda.synchronize()
        assert torch.equal(xg, xe), k
    assert rel_l2(xe.cpu().numpy(), v.cpu().numpy()) < 0.5
    pe.close()
    pg.close()


# ---- several problems: SART's adjoint on the separable tiles ----
def _mixed_class_problems(seed=31):
    """Problems of every z-cell class of the separable tiles (nz up to 64,
    128, 192, 256), one with indicator holes, one whose views include tilted
    ones and one z-driven problem (both of those stay on the brick kernel)."""
    from ct2 import configs3D, phantoms
    from ct2.operator import ConeBeamOperator
    rng = np.random.default_rng(seed)
    specs = [((14, 12, 20), 9, 30, 36, False, False), ((12, 10, 70), 12, 24, 100, True, False),
             ((10, 11, 130), 10, 24, 180, False, False), ((9, 8, 200), 11, 22, 270, False, False),
             ((8, 8, 256), 7, 20, 340, False, False), ((16, 14, 24), 10, 30, 40, False, True)]
    ops, vols, sinos, sep = [], [], [], []
    for i, (shape, ns, nu, nv, holes, tilt) in enumerate(specs):
        fv = configs3D.frame_vectors("circular_conebeam", {"radius": 300.0,
                                                          "source_to_detector": 600.0},
                                     ns, 2 * np.pi, 0.4 * i, 0).astype(np.float64)
        if tilt:
            for p in range(0, ns, 4):
                eu, ev = fv[p, 9:12].copy(), fv[p, 12:15].copy()
                a = np.deg2rad(2.0)
                fv[p, 12:15] = np.cos(a) * ev + np.sin(a) * eu
                fv[p, 9:12] = np.cos(a) * eu - np.sin(a) * ev
        ind = (rng.random((ns, nu, nv)) > 0.05).astype(np.int8) if holes else None
        hw = np.array(shape) / 2.0
        op = ConeBeamOperator(fv.astype(np.float32), ns, nu, nv, 1.6, 1.6, -nu * 0.8, -nv * 0.8,
                              shape, (1.0, 1.0, 1.0), tuple(-hw), arith="fast", indicator=ind)
        n = max(shape)
        cube = phantoms.random_ellipsoids(n, 0.45 * n, 6, seed + i)[:shape[0], :shape[1], :shape[2]]
        v = torch.from_numpy(np.ascontiguousarray(cube)).cuda()
        ops.append(op)
        vols.append(v)
        sinos.append(op.forward(v))
        sep.append(not tilt)
    fv = configs3D.frame_vectors("tomosynthesis", {"radius": 650.0, "center_to_detector": 20.0},
                                 11, np.deg2rad(50), np.deg2rad(65), 1)
    op = ConeBeamOperator(fv, 11, 60, 50, 1.0, 1.0, -30.0, -25.0, (48, 40, 10), (0.8, 0.8, 2.0),
                          (-19.2, -16.0, 0.0), axis="z", arith="fast")
    v = torch.from_numpy(rng.random((48, 40, 10)).astype(np.float32)).cuda()
    ops.append(op)
    vols.append(v)
    sinos.append(op.forward(v))
    sep.append(False)
    return ops, vols, sinos, sep


def test_batch_sart_separable_tiles_equal_sart_step():
    """Problems on the separable tiles give ct2_sart_step's updates bit for
    bit, in every z-cell class; the problems left to the brick kernel agree
    to summation order; the plan's graph replays equal eager calls."""
    from ct2.batch import BatchPlan
    from ct2.solvers import capture
    ops, vols, sinos, sep = _mixed_class_problems()
    nsub, lam, niter = 3, 0.9, 2
    xb = [torch.zeros_like(v) for v in vols]
    plan = BatchPlan(ops, sinos, xb, nsub=nsub, tune=True)
    plan.sart(lam, niter)
    for i, op in enumerate(ops):
        x = _sart_steps(op, sinos[i], nsub, lam, niter)
        if sep[i]:
            assert torch.equal(xb[i], x), (i, rel_l2(xb[i].cpu().numpy(), x.cpu().numpy()))
        else:
            assert rel_l2(xb[i].cpu().numpy(), x.cpu().numpy()) < 1e-5, i
        assert rel_l2(x.cpu().numpy(), vols[i].cpu().numpy()) < 0.9
    # graph replays with the relaxation read from the device
    from ct2 import _dev
    from ct2._lib import call
    xg = [torch.zeros_like(v) for v in vols]
    xe = [torch.zeros_like(v) for v in vols]
    pg = BatchPlan(ops, sinos, xg, nsub=nsub, tune=False)
    pe = BatchPlan(ops, sinos, xe, nsub=nsub, tune=False)
    lamd = torch.full((1,), 1.0, dtype=torch.float32, device="cuda")

    def step(p):
        call("ct2_batch_sart_dev", p._h, _dev.ptr(lamd), 1, _dev.stream_handle())

    step(pg)
    g = capture(lambda: step(pg))
    step(pe)
    for val in (0.7, 0.5):
        lamd.fill_(val)
        g.replay()
        step(pe)
        torch.cuda.synchronize()
        for a, b in zip(xg, xe):
            assert torch.equal(a, b)
    for p in (plan, pg, pe):
        p.close()


def test_batch_sart_separable_tiles_vs_bricks(monkeypatch):
    """CT2_PLAN_SEP=0 keeps every problem on the brick kernel: the two plans
    agree to FP32 summation order, and the separable one holds its tables and
    correction volumes."""
    from ct2.batch import BatchPlan
    ops, vols, sinos, sep = _mixed_class_problems(seed=32)
    xs = [torch.zeros_like(v) for v in vols]
    xb = [torch.zeros_like(v) for v in vols]
    ps = BatchPlan(ops, sinos, xs, nsub=4, tune=False)
    monkeypatch.setenv("CT2_PLAN_SEP", "0")
    pb = BatchPlan(ops, sinos, xb, nsub=4, tune=False)
    ps.sart(1.0, 2)
    pb.sart(1.0, 2)
    for i in range(len(ops)):
        e = rel_l2(xs[i].cpu().numpy(), xb[i].cpu().numpy())
        assert e < 1e-5, (i, e)
    assert pb.info()["bytes"] < ps.info()["bytes"]
    ps.close()
    pb.close()
