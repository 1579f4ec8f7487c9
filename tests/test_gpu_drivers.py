"""Iteration drivers: EM + TV descent captured in a HIP graph, ASD-POCS with
checkpoint/resume (bit for bit) and the traffic-file stop, FDK -> EM-TV, and
both loops against their compositions from the FP32 oracle
(oracle/drivers.py: jawPhantomCCB.py:191-327 and the EM+TV pattern)."""
import numpy as np
import pytest
import torch

from cases import cone_case, rel_l2

pytestmark = pytest.mark.gpu


def small_problem(n=48, ns=40, nd=72, seed=5, photons=None):
    from ct2 import configs3D, phantoms
    from ct2.operator import ConeBeamOperator
    R, SDD, dx, du = 500.0, 1000.0, 1.0, 1.6
    fv = configs3D.frame_vectors("circular_conebeam", {"radius": R, "source_to_detector": SDD},
                                 ns, 2 * np.pi, 0.0, 0)
    hw = n * dx / 2
    op = ConeBeamOperator(fv, ns, nd, nd, du, du, -nd * du / 2, -nd * du / 2, (n, n, n),
                          (dx, dx, dx), (-hw, -hw, -hw), arith="fast")
    truth = torch.from_numpy(np.ascontiguousarray(
        phantoms.random_ellipsoids(n, hw, 12, seed) * 0.02, np.float32)).cuda()
    b = op.forward(truth)
    if photons:
        rng = np.random.default_rng(seed + 1)
        p = b.cpu().numpy().astype(np.float64)
        I = rng.poisson(photons * np.exp(-p))
        b = torch.from_numpy((-np.log(np.maximum(I, 1) / photons)).astype(np.float32)).cuda()
    return op, truth, b


def test_distance_kernel():
    from ct2.solvers import distance
    rng = np.random.default_rng(1)
    a = rng.standard_normal(300001).astype(np.float32)
    b = rng.standard_normal(300001).astype(np.float32)
    m = (rng.random(300001) < 0.7).astype(np.int8)
    ta, tb, tm = (torch.from_numpy(v).cuda() for v in (a, b, m))
    got = distance(ta, tb, tm).item()
    ref = np.sqrt(np.sum(m * (a.astype(np.float64) - b) ** 2))
    assert abs(got - ref) <= 1e-12 * ref
    assert abs(distance(ta).item() - np.linalg.norm(a.astype(np.float64))) <= 1e-12 * ref


def test_em_tv_graph_replays_match_eager():
    """One captured EM+TV iteration replayed; each replay (with eager work and
    synchronisations in between) matches the eager iteration."""
    from ct2.solvers import EmTv, capture
    op, truth, b = small_problem(photons=1e5)
    x0 = torch.full(op.shape, 0.01, device="cuda")
    xe = x0.clone()
    eager = EmTv(op, b, xe, n_tv=5, eps=1e-8)
    ref = []
    for _ in range(5):
        eager.iterate()
        ref.append(xe.clone())
    xg = x0.clone()
    st = EmTv(op, b, xg, n_tv=5, eps=1e-8)
    st.iterate()
    assert rel_l2(xg.cpu().numpy(), ref[0].cpu().numpy()) < 1e-5
    g = capture(st.iterate)
    # capture records without running; state is still after iteration 1
    assert rel_l2(xg.cpu().numpy(), ref[0].cpu().numpy()) < 1e-5
    for k in range(1, 5):
        g.replay()
        torch.cuda.synchronize()
        junk = torch.randn(1000, device="cuda").sum().item()   # eager work between replays
        assert np.isfinite(junk)
        err = rel_l2(xg.cpu().numpy(), ref[k].cpu().numpy())
        assert err < 1e-4, (k, err)
    assert float(st.gf.item()) > 0


def test_em_tv_reduces_error():
    from ct2.solvers import em_tv
    op, truth, b = small_problem(photons=1e5)
    x = torch.full(op.shape, 0.01, device="cuda")
    e0 = rel_l2(x.cpu().numpy(), truth.cpu().numpy())
    em_tv(op, b, x, n_em=10, n_tv=5, eps=1e-8, graph=True)
    e1 = rel_l2(x.cpu().numpy(), truth.cpu().numpy())
    assert float(x.min()) >= 0.0
    assert e1 < 0.6 * e0, (e0, e1)


def _jaw_objects(n=32, ns=24, nu=48, nv=48):
    from ct2 import tomo3D
    ind, fv, geo, vol, g2 = cone_case(n=n, ns=ns, nu=nu, nv=nv, holes=False)
    sino = tomo3D.sinogram3D('circular_conebeam', {"radius": 500.0, "source_to_detector": 1000.0},
                             (ns, nu, nv), 2 * np.pi, nu * 1.5, nv * 1.5, 0.0, -nu * 0.75,
                             -nv * 0.75, 0)
    img = tomo3D.image3D((n, n, n), n * 1.0, n * 1.0, n * 1.0, -n / 2, -n / 2, -n / 2)
    truth = img.duplicate()
    truth.mat = np.ascontiguousarray(vol, np.float32)
    truth.to_device()
    sino.to_device()
    img.to_device()
    truth.project_to(sino)
    return img, sino, truth


def test_asd_pocs_resume_continues_trajectory(tmp_path):
    from ct2.solvers import asd_pocs
    img, sino, truth = _jaw_objects()
    work = sino.duplicate()
    a = img.duplicate()
    rec_a = asd_pocs(a, sino, work, max_iter=4, ngradmax=5, truth=truth)
    b = img.duplicate()
    rec_b1 = asd_pocs(b, sino, work, max_iter=2, ngradmax=5, truth=truth,
                      checkpoint_dir=str(tmp_path))
    c = img.duplicate()
    c.mat = torch.zeros_like(c.mat)
    rec_b2 = asd_pocs(c, sino, work, max_iter=4, ngradmax=5, truth=truth,
                      checkpoint_dir=str(tmp_path), resume=True)
    assert [r[0] for r in rec_b1 + rec_b2] == [1, 2, 3, 4]
    # the state files hold the image and gf, reg_fac, meps exactly (repr),
    # and every device step is deterministic: the same trajectory bit for bit
    assert np.array_equal(c.mat.cpu().numpy(), a.mat.cpu().numpy())
    # the reference's state files and log
    from ct2 import Utils
    assert (tmp_path / "iternum.txt").read_text() == "4"
    log = Utils.read_derr(str(tmp_path / "derrEvol.txt"))
    np.testing.assert_array_equal(log, np.array(rec_b1 + rec_b2, np.float64))
    assert rec_a == rec_b1 + rec_b2
    # data distance falls
    assert rec_a[-1][1] < rec_a[0][1]


def test_asd_pocs_traffic_file_stop(tmp_path):
    from ct2.solvers import asd_pocs
    img, sino, truth = _jaw_objects(n=24, ns=12, nu=36, nv=36)
    work = sino.duplicate()
    stop = tmp_path / "trafficC.txt"
    stop.write_text("stop\n")
    rec = asd_pocs(img, sino, work, max_iter=10, ngradmax=3, traffic_file=str(stop))
    # the stop is read after the first ART sweep; no record is written then
    assert rec == []


def test_fdk_then_em_tv():
    from ct2.operator import ConeBeamOperator
    from ct2.solvers import fdk_em_tv
    img, sino, truth = _jaw_objects(n=40, ns=60, nu=64, nv=64)
    op = ConeBeamOperator.from_objects(sino, img, arith="fast")
    x = fdk_em_tv(sino, img, op, n_em=5, n_tv=5, graph=True)
    assert x.shape == (40, 40, 40) and bool(torch.isfinite(x).all())
    w = sino.duplicate()
    r0 = sino.dist_to(w)   # w == sino
    assert r0 == 0.0
    w.mat = op.forward(x)
    assert sino.dist_to(w) < 0.5 * float(torch.linalg.vector_norm(sino.mat.double()))


def test_asd_pocs_graph_matches_eager_and_resumes(tmp_path):
    """The graph-captured ASD-POCS iteration (SART data step, device-side gf /
    reg_fac / meps schedule) replays like the eager iteration, logs the
    reference's records, reduces the data distance, and a run cut by a
    checkpoint and resumed from the reference's state files continues the
    same trajectory."""
    from ct2 import Utils
    from ct2.solvers import AsdPocs
    op, truth, b = small_problem(n=40, ns=36, nd=64)
    x0 = torch.zeros(op.shape, device="cuda")
    kw = dict(nsub=6, ngrad=5, iter_eps=3, max_iter=20, truth=truth)
    xe = x0.clone()
    eager = AsdPocs(op, b, xe, **kw)
    rec_e, _ = eager.run(6, graph=False)
    xg = x0.clone()
    gr = AsdPocs(op, b, xg, **kw)
    rec_g, stopped = gr.run(6, graph=True, checkpoint_dir=str(tmp_path), save_every=2)
    assert not stopped and rec_g.shape == (6, 4)
    assert rel_l2(xg.cpu().numpy(), xe.cpu().numpy()) < 1e-5
    np.testing.assert_allclose(rec_g, rec_e, rtol=1e-5)
    assert list(rec_g[:, 0]) == [1, 2, 3, 4, 5, 6]
    assert rec_g[-1, 1] < 0.5 * rec_g[0, 1] and rec_g[-1, 3] < rec_g[0, 3]
    np.testing.assert_allclose(Utils.read_derr(str(tmp_path / "derrEvol.txt")), rec_g, rtol=1e-12)
    it, gf, reg_fac, meps = Utils.read_state(str(tmp_path))
    assert it == 6 and reg_fac == pytest.approx(0.995 ** 6, rel=1e-6)
    assert meps == pytest.approx(rec_g[2, 1], rel=1e-12)       # distance at iter_eps
    # cut after 3 iterations, resume from the files, 3 more
    d2 = tmp_path / "cut"
    xa = x0.clone()
    a = AsdPocs(op, b, xa, **kw)
    a.run(3, graph=True, checkpoint_dir=str(d2), save_every=3)
    xb = torch.full(op.shape, 7.0, device="cuda")
    c = AsdPocs(op, b, xb, **kw)
    c.resume(str(d2))
    rec_c, _ = c.run(3, graph=True, checkpoint_dir=str(d2), save_every=3)
    assert list(rec_c[:, 0]) == [4, 5, 6]
    assert rel_l2(xb.cpu().numpy(), xg.cpu().numpy()) < 1e-5
    assert Utils.read_derr(str(d2 / "derrEvol.txt")).shape == (6, 4)


def test_asd_pocs_graph_traffic_stop(tmp_path):
    from ct2.solvers import AsdPocs
    op, truth, b = small_problem(n=32, ns=24, nd=48)
    x = torch.zeros(op.shape, device="cuda")
    stop = tmp_path / "traffic.txt"
    stop.write_text("stop\n")
    st = AsdPocs(op, b, x, nsub=4, ngrad=3, max_iter=50)
    rec, stopped = st.run(20, graph=True, save_every=4, traffic_file=str(stop))
    assert stopped and rec.shape[0] == 4


def _np(a):
    return a.cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)


def test_asd_pocs_matches_oracle_composition():
    """solvers.asd_pocs on device objects (ART, TV value and gradient, the
    forward, the distances) against the loop composed from the FP32 oracle,
    24^3, 3 iterations of one ART sweep and 5 TV steps.  The records agree to
    1e-6 and the image to 1e-5 relative L2, the FP32-mode bar (measured
    2.5e-8 and 7.2e-6: the FP64 sums of the distances run in another order,
    and the TV descent's normalised steps carry a last-bit difference of gf
    into the image)."""
    import oracle
    from oracle import drivers as odr
    from ct2.solvers import asd_pocs
    img, sino, truth = _jaw_objects(n=24, ns=12, nu=36, nv=36)
    geo = (sino.ns, sino.nu, sino.nv, sino.du, sino.dv, sino.u0, sino.v0)
    g2 = (img.dx, img.dy, img.dz, img.x0, img.y0, img.z0, img.nx, img.ny, img.nz)
    x0 = _np(img.mat).copy()
    work = sino.duplicate()
    x = img.duplicate()
    rec = asd_pocs(x, sino, work, max_iter=3, ngradmax=5, truth=truth)
    o32 = oracle.restated_raw("f32")
    xo, reco, _ = odr.asd_pocs(o32, x0, _np(sino.mat), _np(sino.indicator).astype(np.int8),
                               _np(sino.frame_vectors), geo, g2, max_iter=3, ngradmax=5,
                               truth=_np(truth.mat))
    got = _np(x.mat)
    e = rel_l2(got, xo)
    print(f"ASD-POCS vs oracle composition: image rel L2 {e:.2e}; records {rec} / {reco}")
    assert e <= 1e-5, e
    assert [r[0] for r in rec] == [r[0] for r in reco] == [1, 2, 3]
    for r, ro in zip(rec, reco):
        for k in (1, 2, 3):
            assert abs(r[k] - ro[k]) <= 1e-6 * abs(ro[k]), (r, ro)


def test_em_tv_matches_oracle_composition():
    """EmTv (EM update, positivity, normalised TV descent with the ASD-POCS
    step rule) with the exact-arithmetic operator against the same loop
    composed from the oracle, 24^3, noisy data with indicator holes, 3
    iterations.  The device adjoint sums in FP64 where the reference's sums
    in FP32, so the bar is the FP64 composition: the device loop lies no
    farther from it than twice the FP32 composition does, or 1e-5."""
    import oracle
    from oracle import drivers as odr
    from ct2.operator import ConeBeamOperator
    from ct2.solvers import EmTv
    ind, fv, geo, vol, g2 = cone_case(n=24, ns=12, nu=36, nv=36, holes=True)
    ns, nu, nv, du, dv, u0, v0 = geo
    dx, dy, dz, x0, y0, z0, nx, ny, nz = g2
    o32 = oracle.restated_raw("f32")
    truth = (np.asarray(vol, np.float32) + 0.05).astype(np.float32)
    rng = np.random.default_rng(11)
    b = o32.sinoproj(0, 0, ind, fv, *geo, truth, *g2)
    b = np.maximum(b * (1 + 0.02 * rng.standard_normal(b.shape)), 0).astype(np.float32)
    op = ConeBeamOperator(fv, ns, nu, nv, du, dv, u0, v0, (nx, ny, nz), (dx, dy, dz),
                          (x0, y0, z0), arith="exact", indicator=ind)
    start = np.full((nx, ny, nz), 0.1, np.float32)
    x = torch.from_numpy(start.copy()).cuda()
    st = EmTv(op, torch.from_numpy(b).cuda(), x, n_tv=5, eps=1e-8)
    for _ in range(3):
        st.iterate()
    xo, hist = odr.em_tv(o32, start, b, ind, fv, geo, g2, n_em=3, n_tv=5, eps=1e-8)
    x64, _ = odr.em_tv(oracle.restated_raw("f64"), start, b, ind, fv, geo, g2, n_em=3,
                       n_tv=5, eps=1e-8)
    e64, e32 = rel_l2(_np(x), x64), rel_l2(xo, x64)
    print(f"EM+TV: device vs FP64 composition {e64:.2e}, FP32 composition vs FP64 {e32:.2e}, "
          f"device vs FP32 {rel_l2(_np(x), xo):.2e}; gf {float(st.gf.item())} / {hist[-1][2]}")
    assert e64 <= max(2 * e32, 1e-5), (e64, e32)
    assert abs(float(st.gf.item()) - hist[-1][2] * (0.98 if hist[-1][1] >= 0.99 * hist[-1][0]
                                                    else 1.0)) <= 1e-4 * hist[-1][2]


def test_asd_pocs_sart_graph_matches_oracle_composition():
    """AsdPocs (ordered-subset SART data step, schedule on the device, three
    iterations through the captured HIP graph: one eager, two replays)
    against the same loop composed from the oracle (oracle/drivers.py
    asd_pocs_sart), 24^3 with indicator holes, exact arithmetic, 4 subsets and
    5 TV steps.  The device SART sums its adjoint in FP64 where the FP32
    composition sums in FP32, so the bar is the FP64 composition: the device
    trajectory lies no farther from it than twice the FP32 composition does,
    or 1e-5; the data-distance records agree to 1e-5."""
    import oracle
    from oracle import drivers as odr
    from ct2.operator import ConeBeamOperator
    from ct2.solvers import AsdPocs
    ind, fv, geo, vol, g2 = cone_case(n=24, ns=12, nu=36, nv=36, holes=True, seed=2)
    ns, nu, nv, du, dv, u0, v0 = geo
    dx, dy, dz, x0, y0, z0, nx, ny, nz = g2
    o32 = oracle.restated_raw("f32")
    truth = np.asarray(vol, np.float32)
    b = o32.sinoproj(0, 0, ind, fv, *geo, truth, *g2)
    op = ConeBeamOperator(fv, ns, nu, nv, du, dv, u0, v0, (nx, ny, nz), (dx, dy, dz),
                          (x0, y0, z0), arith="exact", indicator=ind)
    x = torch.zeros((nx, ny, nz), dtype=torch.float32, device="cuda")
    st = AsdPocs(op, torch.from_numpy(b).cuda(), x, nsub=4, ngrad=5)
    rec, stopped = st.run(3, graph=True)
    assert not stopped and [int(r[0]) for r in rec] == [1, 2, 3]
    start = np.zeros((nx, ny, nz), np.float32)
    kw = dict(max_iter=3, nsub=4, ngrad=5)
    x32, rec32, _ = odr.asd_pocs_sart(o32, start, b, ind, fv, geo, g2, **kw)
    x64, _, _ = odr.asd_pocs_sart(oracle.restated_raw("f64"), start, b, ind, fv, geo, g2, **kw)
    got = _np(x)
    e64, e32 = rel_l2(got, x64), rel_l2(x32, x64)
    print(f"AsdPocs: device vs FP64 composition {e64:.2e}, FP32 composition vs FP64 {e32:.2e}, "
          f"device vs FP32 {rel_l2(got, x32):.2e}")
    assert e64 <= max(2 * e32, 1e-5), (e64, e32)
    for r, ro in zip(rec, rec32):
        assert abs(r[1] - ro[1]) <= 1e-5 * abs(ro[1]), (r, ro)
