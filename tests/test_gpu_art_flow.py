"""ART across the GPU: the dataflow schedule (k_art_flow: persistent
wavefronts over detector columns, each ray waiting only for the earlier rays
that can share a voxel with it) against the single-workgroup level schedule
(k_art_views, CT2_ART_FLOW=0).  Both run every pair of rays that share a
voxel in the reference's (view, u, v) order with the same per-ray arithmetic,
so the volumes agree bit for bit.

Cases: the C1 geometry (64^3, 90 views, 96x96 detector) over two sweeps,
indicator holes with linear and nearest sampling, a view range that starts
mid-scan, and the z-driven tomosynthesis ART with its second volume
(artstepz2vol).  The error path: with the give-up time of a wait set to one
tick from the host, waits give up, every wavefront leaves, and the call
raises instead of hanging.
"""
import os

import numpy as np
import pytest

from cases import cone_case, tomo_case

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _level_and_flow(fn):
    os.environ["CT2_ART_FLOW"] = "0"
    try:
        level = fn()
    finally:
        del os.environ["CT2_ART_FLOW"]
    flow = fn()
    return level, flow


def _c1():
    with np.load(os.path.join(HERE, "golden", "c1_reference.npz"), allow_pickle=False) as z:
        c1 = {k: z[k] for k in z.files}
    g = c1["geo"]
    geo = (int(g[0]), int(g[1]), int(g[2]), *[float(v) for v in g[3:]])
    gg = c1["grid"]
    g2 = (*[float(v) for v in gg[:6]], int(gg[6]), int(gg[7]), int(gg[8]))
    return c1, geo, g2


def test_c1_two_sweeps_bit_for_bit():
    from ct2 import view_proj3D as vp
    c1, geo, g2 = _c1()
    fv, sino = c1["frame_vectors"], c1["sino"]

    def run():
        x = np.zeros_like(c1["vol"])
        for _ in range(2):
            x = vp.artstep_linear(0.5, sino, None, fv, *geo, x, *g2)
        return x

    level, flow = _level_and_flow(run)
    assert np.abs(flow).max() > 0
    assert np.array_equal(level, flow)


@pytest.mark.parametrize("interp", ["linear", "nearest"])
def test_holes_bit_for_bit(interp):
    from ct2 import view_proj3D as vp
    ind, fv, geo, vol, g2 = cone_case(n=32, ns=20, nu=48, nv=44, holes=True, seed=4)
    sino = vp.sinoproj_linear(None, fv, *geo, vol, *g2)
    art = getattr(vp, f"artstep_{interp}")
    x0 = np.full_like(vol, 0.02)
    level, flow = _level_and_flow(lambda: art(0.8, sino, ind, fv, *geo, x0, *g2))
    assert np.array_equal(level, flow)


def test_view_range_bit_for_bit():
    import ctypes as C
    import torch
    from ct2 import _dev
    from ct2._lib import call, ct2_desc
    ind, fv, geo, vol, g2 = cone_case(n=24, ns=16, nu=36, nv=36, holes=True, seed=5)
    ns, nu, nv, du, dv, u0, v0 = geo
    dx, dy, dz, x0, y0, z0, nx, ny, nz = g2
    d = ct2_desc()
    call("ct2_desc_init", C.byref(d), ns, nu, nv, du, dv, u0, v0, nx, ny, nz, dx, dy, dz, x0,
         y0, z0)
    d.view0, d.nviews = 5, 9
    rng = np.random.default_rng(6)
    sino = torch.from_numpy(rng.random((ns, nu, nv)).astype(np.float32)).cuda()
    indd, fvd = torch.from_numpy(ind).cuda(), torch.from_numpy(fv).cuda()

    def run():
        x = torch.zeros((nx, ny, nz), dtype=torch.float32, device="cuda")
        call("ct2_art_step", C.byref(d), 0.6, _dev.ptr(sino), _dev.ptr(indd), _dev.ptr(fvd),
             _dev.ptr(x), _dev.stream_handle())
        torch.cuda.synchronize()
        return x.cpu().numpy()

    level, flow = _level_and_flow(run)
    assert np.abs(flow).max() > 0
    assert np.array_equal(level, flow)


def test_tomo_two_volume_bit_for_bit():
    from ct2 import view_proj3D_tomo as vpt
    ind, fv, geo, vol, g2 = tomo_case(ns=5)
    sino = vpt.sinoprojz(ind, fv, *geo, vol, *g2)
    rng = np.random.default_rng(3)
    va = rng.random((20, 16, 4)).astype(np.float32)
    ga = (1.6, 1.6, 3.0, -16.0, -12.8, 16.0, 20, 16, 4)
    level, flow = _level_and_flow(lambda: vpt.artstepz(0.7, sino, ind, fv, *geo,
                                                      np.zeros_like(vol), *g2))
    assert np.array_equal(level, flow)
    level, flow = _level_and_flow(lambda: vpt.artstepz2vol(0.7, sino, ind, fv, *geo, vol, *g2,
                                                          va, *ga))
    assert np.array_equal(level[0], flow[0]) and np.array_equal(level[1], flow[1])


def test_wait_give_up_raises():
    from ct2 import view_proj3D as vp
    from ct2._lib import CT2Error
    c1, geo, g2 = _c1()
    os.environ["CT2_ART_WAIT_TICKS"] = "1"
    try:
        with pytest.raises(CT2Error, match="gave up"):
            vp.artstep_linear(0.5, c1["sino"], None, c1["frame_vectors"], *geo,
                              np.zeros_like(c1["vol"]), *g2)
    finally:
        del os.environ["CT2_ART_WAIT_TICKS"]
    # the next call, with the normal limit, runs to the end
    x = vp.artstep_linear(0.5, c1["sino"], None, c1["frame_vectors"], *geo,
                          np.zeros_like(c1["vol"]), *g2)
    assert np.isfinite(x).all() and np.abs(x).max() > 0
