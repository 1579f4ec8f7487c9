"""Headline benchmark: giga voxel-ray updates per second of one Joseph
forward projection plus one exact-adjoint backprojection of a 256^3 volume,
360 views, 512x512 detector, with 16-bit (FP16) volume storage and FP32
accumulation ("mixed"), next to the plain FP32 path, on one MI355X.

Workload (one step = fwd + bp over all views, inputs resident in HBM):
  volume  256^3, 0.8 mm voxels, 40 seeded random ellipsoids (seed 20260101)
  scan    circular cone beam, 360 views over 360 deg, source-isocentre 500 mm,
          source-detector 1000 mm, 512x512 detector at 0.8 mm
  updates N = 2 * 360 * 512 * 512 * 256 = 4.832e10 per step
Also reported: the FP32-fast path (fast arithmetic, f32 volume storage) and
the plain FP32 path (reference operation order), the relative L2 error of the
mixed and FP32 paths against the FP64 oracle on a subset of views, the
dominant kernel's roofline (its binding roof, VALU issue: the executed
VALU instructions per launch from the committed rocprofv3 counter summary
over the live duration, against the issue rates ct2_calib_issue measures on
the same GPU for the kernel's FP64/other instruction mix; beside it the
compulsory HBM bytes and the measured HBM traffic against 8 TB/s), and the
reference path (the reference Fortran compiled from its sources, 1 thread)
timed on a bounded sample of the same workload.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-check]

The oracle (CPU restatement and compiled reference) is used only in the CPU
leg (cpu_leg), as the timed baseline and as the checker of the GPU outputs.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "ct-2_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = ("Gvox-ray updates/s (fwd+bp, 256³, 360 views), mixed vs FP32 at "
          "≤1e-4 rel err")
HBM_PEAK_GBS = 8000.0
N, NS, ND = 256, 360, 512
VOX, R, SDD, PITCH = 0.8, 500.0, 1000.0, 0.8


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def geometry():
    import numpy as np
    from ct2 import configs3D
    fv = configs3D.frame_vectors("circular_conebeam", {"radius": R, "source_to_detector": SDD},
                                 NS, 2 * np.pi, 0.0, 0)
    hw = N * VOX / 2
    geo = (NS, ND, ND, PITCH, PITCH, -ND * PITCH / 2, -ND * PITCH / 2)
    grid = (VOX, VOX, VOX, -hw, -hw, -hw, N, N, N)
    return fv, geo, grid


def phantom():
    from ct2 import phantoms
    return phantoms.random_ellipsoids(N, N * VOX / 2, 40, 20260101)


def make_op(fv, geo, grid, arith):
    from ct2.operator import ConeBeamOperator
    ns, nu, nv, du, dv, u0, v0 = geo
    dx, dy, dz, x0, y0, z0, nx, ny, nz = grid
    return ConeBeamOperator(fv, ns, nu, nv, du, dv, u0, v0, (nx, ny, nz), (dx, dy, dz),
                            (x0, y0, z0), arith=arith)


def to_storage(vol32, vol16, stochastic, seed=20260107):
    """FP32 volume -> FP16 storage through the library's conversion kernel;
    a (*shape, 2) target receives the f16 pair storage."""
    from ct2.operator import to_storage as convert
    return convert(vol32, vol16.dtype, "stochastic" if stochastic else "nearest", seed, out=vol16,
                   pairs=vol16.dim() == 4)


def time_path(op, vol32, vol_store, sino, data, recon, steps, warmup):
    """Time `steps` steps of [FP32 -> storage conversion (mixed only), forward
    of the volume into `sino`, adjoint of the measured data `data` (Poisson-
    noised, so no ray is zero and the adjoint skips none)]; per-kernel
    durations from HIP events on the stream the kernels run on (torch's
    current stream)."""
    import torch

    def step(ev=None):
        if ev: ev[0].record()
        if vol_store is not None:
            to_storage(vol32, vol_store, stochastic=True)
        v = vol_store if vol_store is not None else vol32
        if ev: ev[1].record()
        op.forward(v, out=sino)
        if ev: ev[2].record()
        op.adjoint(data, out=recon)
        if ev: ev[3].record()

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(steps)]
    t0 = time.perf_counter()
    for s in range(steps):
        step(ev[s])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t_cvt = sum(e[0].elapsed_time(e[1]) for e in ev) / steps
    t_fwd = sum(e[1].elapsed_time(e[2]) for e in ev) / steps
    t_bp = sum(e[2].elapsed_time(e[3]) for e in ev) / steps
    return wall, t_cvt, t_fwd, t_bp


def noisy_data(op, vol32, photons=1e5, seed=20260102):
    """Poisson-noised projections p_hat = -ln(max(I,1)/N), I ~ Poisson(N e^-p),
    of the phantom (attenuation 0.02/mm per unit density)."""
    import torch
    p = op.forward(vol32 * 0.02)
    g = torch.Generator(device="cuda").manual_seed(seed)
    counts = torch.poisson(photons * torch.exp(-p.double()), generator=g)
    return (-torch.log(torch.clamp(counts, min=1.0) / photons)).float().contiguous()


def check_views(ns, count=4):
    return list(range(0, ns, ns // count))[:count]


def gpu_precision_outputs(op_mixed, op_fp32, vol32, vol16, sub):
    """GPU outputs on the views `sub` that the CPU leg checks against the FP64
    oracle: the forward in each storage mode and the adjoint (both
    arithmetics) of the FP32 forward restricted to those views."""
    import torch
    out = {}
    s = torch.empty(op_mixed.sino_shape, dtype=torch.float32, device=op_mixed.device)
    to_storage(vol32, vol16, stochastic=True)
    op_mixed.forward(vol16, out=s)
    out["fwd_mixed"] = s[sub].cpu().numpy()
    to_storage(vol32, vol16, stochastic=False)
    op_mixed.forward(vol16, out=s)
    out["fwd_mixed_round_nearest"] = s[sub].cpu().numpy()
    op_fp32.forward(vol32, out=s)
    out["fwd_fp32"] = s[sub].cpu().numpy()
    y = torch.zeros_like(s)
    y[sub] = s[sub]
    out["y"] = out["fwd_fp32"]
    out["bp_mixed"] = op_mixed.adjoint(y).cpu().numpy()
    out["bp_fp32"] = op_fp32.adjoint(y).cpu().numpy()
    return out


def cpu_leg(fv, geo, grid, vol_np, views=8, gpu_out=None, sub=None):
    """The CPU leg.  (1) The reference path timed on the host:
    sinoproj_linear + sinobackproject_linear of the reference Fortran
    (oracle/_ref, compiled from the reference's own sources) on `views` views
    of the workload, 1 thread; the FP32 C restatement (kind "port") stands in
    when the compiled reference is not present.  (2) As the checker: the FP64
    oracle on the views `sub`, against which the GPU outputs `gpu_out` are
    scored (relative L2)."""
    import numpy as np
    import oracle
    ns, nu, nv = geo[:3]
    tsub = list(range(0, ns, ns // views))[:views]
    fv_sub = np.ascontiguousarray(fv[tsub])
    geo_sub = (len(tsub),) + tuple(geo[1:])
    if oracle.ref_available():
        mods, kind = oracle.reference(), "reference"
    else:
        mods, kind = oracle.restated("f32"), "port"
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    t0 = time.perf_counter()
    sref = mods.view_proj3D.sinoproj_linear(None, fv_sub, *geo_sub, vol_np, *grid)
    t1 = time.perf_counter()
    mods.view_proj3D.sinobackproject_linear(sref, None, fv_sub, *geo_sub, *grid)
    t2 = time.perf_counter()
    nupd = 2 * len(tsub) * nu * nv * N
    base = {"value": round(nupd / (t2 - t0) / 1e9, 6), "unit": "Gvox-ray updates/s",
            "cores": 1, "kind": kind,
            "sample": (f"{len(tsub)} of {ns} views (fwd {t1 - t0:.1f} s + bp {t2 - t1:.1f} s), "
                       f"{N}^3 volume, {nu}x{nv} detector, single thread"
                       + (", reference Fortran compiled by amdflang -O2 from emil_code/"
                          "view_proj3D.f, called through ctypes" if kind == "reference" else
                          ", FP32 C restatement of the reference"))}
    prec = None
    if gpu_out is not None:
        o64 = oracle.restated_raw("f64")
        fv_c = np.ascontiguousarray(fv[sub])
        geo_c = (len(sub),) + tuple(geo[1:])
        ref = o64.sinoproj(oracle.LINEAR, 0, None, fv_c, *geo_c, vol_np, *grid)
        refb = o64.sinobackproject(oracle.LINEAR, 0, gpu_out["y"], None, fv_c, *geo_c, *grid)

        def rel(a, b):
            a = np.asarray(a, np.float64)
            return float(np.linalg.norm(a - b) / np.linalg.norm(b))
        prec = {"views_checked": len(sub),
                "mixed": {"fwd": rel(gpu_out["fwd_mixed"], ref),
                          "bp": rel(gpu_out["bp_mixed"], refb)},
                "mixed_round_nearest_fwd": rel(gpu_out["fwd_mixed_round_nearest"], ref),
                "fp32": {"fwd": rel(gpu_out["fwd_fp32"], ref),
                         "bp": rel(gpu_out["bp_fp32"], refb)}}
    return base, prec


def cpu_multicore(fv, geo, grid, vol_np, max_threads=16):
    """The FP32 C restatement of the reference path on all host cores this
    process may use (at most max_threads): one view per thread, each thread
    running sinoproj_linear + sinobackproject_linear on its own view (the
    ctypes calls release the GIL, and the restatement keeps no shared state).
    The reference Fortran itself is single-threaded and is not run this way."""
    import concurrent.futures as cf
    import numpy as np
    import oracle
    mods = oracle.restated("f32")
    ns, nu, nv = geo[:3]
    threads = max(1, min(max_threads, len(os.sched_getaffinity(0))))
    views = list(range(0, ns, max(1, ns // threads)))[:threads]
    geo1 = (1,) + tuple(geo[1:])

    def one(p):
        fvp = np.ascontiguousarray(fv[p:p + 1])
        s = mods.view_proj3D.sinoproj_linear(None, fvp, *geo1, vol_np, *grid)
        mods.view_proj3D.sinobackproject_linear(s, None, fvp, *geo1, *grid)

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        list(ex.map(one, views))
    t1 = time.perf_counter()
    nupd = 2 * len(views) * nu * nv * N
    return {"value": round(nupd / (t1 - t0) / 1e9, 6), "unit": "Gvox-ray updates/s",
            "cores": threads, "kind": "port",
            "sample": (f"{len(views)} of {ns} views, one per thread ({t1 - t0:.1f} s wall), "
                       f"{N}^3 volume, {nu}x{nv} detector, FP32 C restatement of the "
                       f"reference, {os.cpu_count()} CPUs visible")}


def issue_rates(blocks=2048, iters=20000):
    """Measured VALU issue rates of this GPU, wavefront instructions per
    second, for FP32 FMA, FP64 FMA and 32-bit integer rotates (ct2_calib_issue:
    8 independent chains per thread, 8 wavefronts per SIMD), best of 3."""
    import torch
    from ct2 import _dev
    from ct2._lib import call
    out = torch.empty(blocks * 256, dtype=torch.float32, device="cuda")
    rates = {}
    for kind, name in ((0, "f32_fma"), (1, "f64_fma"), (2, "i32_rotate")):
        best = 1e30
        for rep in range(4):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            call("ct2_calib_issue", kind, blocks, iters, _dev.ptr(out), _dev.stream_handle())
            e1.record()
            torch.cuda.synchronize()
            if rep:   # the first launch warms up
                best = min(best, e0.elapsed_time(e1))
        rates[name] = blocks * 4 * iters * 8 / (best * 1e-3)
    return rates


def read_counters(name="sep_kernels_c2_counters.txt"):
    """Per-launch figures of the dominant kernel (the column-separable
    adjoint) from the committed rocprofv3 counter summary: lines
    "key: value" (valu_insts_per_launch, traffic_bytes_per_launch, ...)."""
    p = os.path.join(ROOT, "profiles", name)
    out = {}
    if not os.path.exists(p):
        return out
    for line in open(p):
        k, sep, v = line.partition(":")
        if sep and k.strip().endswith("_per_launch"):
            try:
                out[k.strip()] = float(v)
            except ValueError:
                pass
    return out


def batch_workload(args):
    """64 independent cone-beam problems (seed 20260104): cubic volumes of
    n ~ U{64..192} voxels of 1 mm, views ~ U{30..180} over 360 deg, detector
    ceil(1.5 n)^2 bins of 1.6 mm, R = 500 mm, SDD = 1000 mm; 20-ellipsoid
    phantoms.  10 SART iterations (10 ordered subsets of interleaved views)
    captured in one HIP graph over a batched plan whose tile and brick were
    picked by the plan's autotuner; the same iteration problem by problem
    through ct2_sart_step is timed for comparison."""
    import numpy as np
    import torch
    from ct2 import configs3D, phantoms
    from ct2.batch import BatchPlan
    from ct2.operator import ConeBeamOperator
    from ct2.solvers import capture
    rng = np.random.default_rng(20260104)
    nprob, nsub, niter, lam = 64, 10, 10, 1.0
    ops, truth, sinos = [], [], []
    log("building 64 problems")
    for p in range(nprob):
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
    x = [torch.zeros_like(v) for v in truth]
    log("planning (autotune, row and column sums)")
    t0 = time.perf_counter()
    plan = BatchPlan(ops, sinos, x, nsub=nsub, tune=True)
    torch.cuda.synchronize()
    t_plan = time.perf_counter() - t0
    info = plan.info()
    upd_iter = plan.updates_per_sart_iteration()
    log(f"  plan {t_plan:.2f} s, tile {info['tile']}, brick {info['brick']}, "
        f"{info['bytes'] / 2**30:.2f} GiB")
    plan.sart(lam, 1)                       # warm-up iteration
    g = capture(lambda: plan.sart(lam, niter))
    for v in x:
        v.zero_()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    g.replay()
    e1.record()
    torch.cuda.synchronize()
    t_graph = e0.elapsed_time(e1) / 1e3
    err = float(np.mean([float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
                         for a, b in zip(x, truth)]))
    # one SART iteration problem by problem (ct2_sart_step per subset)
    log("timing the unbatched iteration")
    xs = [torch.zeros_like(v) for v in truth]
    work = torch.empty(max(2 * s.numel() + 3 * v.numel() for s, v in zip(sinos, truth)),
                       dtype=torch.float32, device="cuda")
    for i, op in enumerate(ops[:2]):        # warm-up (brick tuning of two geometries)
        op.sart_step(sinos[i], xs[i], lam, view0=0, nviews=0, work=work, vstride=nsub)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i, op in enumerate(ops):
        for s in range(nsub):
            op.sart_step(sinos[i], xs[i], lam, view0=s, nviews=0, work=work, vstride=nsub)
    torch.cuda.synchronize()
    t_single = time.perf_counter() - t0
    # the same forward + adjoint work problem by problem (one launch each per
    # subset, no row/column sums): what batching saves in launches and idle CUs
    resid = [torch.empty_like(sv) for sv in sinos]
    corr = [torch.empty_like(v) for v in truth]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i, op in enumerate(ops):
        for s in range(nsub):
            op.forward(xs[i], out=resid[i], view0=s, vstride=nsub)
            op.adjoint(resid[i], out=corr[i], view0=s, vstride=nsub)
    torch.cuda.synchronize()
    t_fb = time.perf_counter() - t0
    value = niter * upd_iter / t_graph / 1e9
    line = {"metric": "Gvox-ray updates/s (batched SART, 64 uneven cone-beam problems, "
                      "10 iterations in one HIP graph)",
            "value": round(value, 3), "unit": "Gvox-ray updates/s", "n_gpus": 1,
            "steps": niter, "warmup": 1, "ms_per_step": round(1e3 * t_graph / niter, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic: seeded random-ellipsoid phantoms and their projections",
            "config": {"workload": "64 problems, n~U{64..192}, views~U{30..180}, detector "
                                   "ceil(1.5n)^2, 10 SART iterations x 10 ordered subsets",
                       "updates_per_iteration": upd_iter},
            "tuned": info, "plan_seconds": round(t_plan, 3),
            "unbatched_fwd_bp_value": round(upd_iter / t_fb / 1e9, 3),
            "unbatched_sart_step_value": round(upd_iter / t_single / 1e9, 3),
            "unbatched_note": "one iteration problem by problem: fwd_bp = one forward and one "
                              "adjoint launch per problem and subset (the batched work); "
                              "sart_step = ct2_sart_step per problem and subset, which also "
                              "recomputes the row and column sums",
            "mean_rel_l2_to_truth_after_10_iterations": round(err, 4)}
    print(json.dumps(line), flush=True)


def recon_workload(args):
    """The 256^3 reconstruction workload: projections of the 40-ellipsoid
    phantom (densities scaled to attenuation 0.02/mm per unit, about water)
    with Poisson noise at 1e5 photons per ray (seed 20260102); FDK, then 20 EM
    iterations each followed by 5 normalised TV-gradient steps (eps 1e-8,
    ASD-POCS step rule), the EM+TV iteration replayed from one HIP graph."""
    import numpy as np
    import torch
    from ct2 import tomo3D
    from ct2.operator import ConeBeamOperator
    from ct2.solvers import EmTv, capture
    from ct2 import FDK
    mu = 0.02
    hw = N * VOX / 2
    sino = tomo3D.sinogram3D("circular_conebeam", {"radius": R, "source_to_detector": SDD},
                             (NS, ND, ND), 2 * np.pi, ND * PITCH, ND * PITCH, 0.0,
                             -ND * PITCH / 2, -ND * PITCH / 2, 0)
    img = tomo3D.image3D((N, N, N), 2 * hw, 2 * hw, 2 * hw, -hw, -hw, -hw)
    op = ConeBeamOperator.from_objects(sino, img, arith="fast")
    truth = torch.from_numpy(phantom()).cuda() * mu
    p = op.forward(truth)
    g = torch.Generator(device="cuda").manual_seed(20260102)
    counts = torch.poisson(1e5 * torch.exp(-p.double()), generator=g)
    b = (-torch.log(torch.clamp(counts, min=1.0) / 1e5)).float().contiguous()
    sino.mat = b
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x = FDK.fdk_reconstruct(b, sino, img).contiguous()
    x.clamp_(min=0.0)
    torch.cuda.synchronize()
    t_fdk = time.perf_counter() - t0
    def err(v):
        return float(torch.linalg.vector_norm(v - truth) / torch.linalg.vector_norm(truth))
    e_fdk = err(x)
    # the reference FDK carries a detector-magnification scale; EM starts
    # from the FDK image rescaled to the data (one scalar, least squares)
    ax = op.forward(x)
    x.mul_(float((ax.double() * b.double()).sum() / (ax.double() * ax.double()).sum()))
    st = EmTv(op, b, x, n_tv=5, eps=1e-8)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st.iterate()                      # eager first iteration (also tunes the kernels)
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t0
    gr = capture(st.iterate)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(19):
        gr.replay()
    torch.cuda.synchronize()
    t_loop = time.perf_counter() - t0
    e_em = err(x)
    resid = float(torch.linalg.vector_norm(op.forward(x) - b) / torch.linalg.vector_norm(b))
    upd = 20 * 2 * op.updates_per_direction()
    line = {"metric": "seconds per reconstruction (FDK + 20 x (EM + 5 TV steps), 256^3, 360 "
                      "views, 512^2 detector)",
            "value": round(t_fdk + t_first + t_loop, 4), "unit": "s", "n_gpus": 1, "steps": 20,
            "warmup": 0, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32",
            "data": "synthetic: 40-ellipsoid phantom x 0.02/mm, Poisson noise 1e5 photons/ray",
            "config": {"workload": "FDK then 20 EM iterations each followed by 5 TV steps"},
            "phases_s": {"fdk": round(t_fdk, 4), "em_tv_first_eager": round(t_first, 4),
                         "em_tv_19_graph_replays": round(t_loop, 4)},
            "em_iteration_ms": round(1e3 * t_loop / 19, 2),
            "em_gupd_per_s": round(19 * 2 * op.updates_per_direction() / t_loop / 1e9, 1),
            "rel_l2_to_truth": {"fdk": round(e_fdk, 4), "fdk_em_tv": round(e_em, 4)},
            "data_residual_rel": round(resid, 5), "updates_em": upd}
    print(json.dumps(line), flush=True)


def art_workload(args):
    """The CPU-runnable workload C1: a 64^3 Shepp-Logan phantom (1 mm
    voxels), circular cone beam of 90 views over 360 deg (R = 500 mm,
    SDD = 1000 mm), 96x96 detector at 1.5 mm; its projection, then 10 ART
    sweeps (artstep_linear, relaxation 0.5) from zero.  Timed on the GPU with
    the dataflow schedule (the default) and with the single-workgroup level
    schedule (CT2_ART_FLOW=0), which must give the same bits; beside them
    the reference Fortran (oracle/_ref, 1 thread; the FP32 C restatement
    when it is not built) on one sweep, the bounded CPU sample."""
    import numpy as np
    import torch
    import oracle
    from ct2 import configs3D, phantoms
    from ct2 import view_proj3D as vp
    ns, nu, nv, du = 90, 96, 96, 1.5
    fv = configs3D.frame_vectors("circular_conebeam", {"radius": 500.0,
                                                      "source_to_detector": 1000.0},
                                 ns, 2 * np.pi, 0.0, 0)
    vol = phantoms.shepp_logan(64, 32.0)
    geo = (ns, nu, nv, du, du, -72.0, -72.0)
    g2 = (1.0, 1.0, 1.0, -32.0, -32.0, -32.0, 64, 64, 64)
    sino_h = vp.sinoproj_linear(None, fv, *geo, vol, *g2)
    sino = torch.from_numpy(sino_h).cuda()
    fvd = torch.from_numpy(np.ascontiguousarray(fv, np.float32)).cuda()

    def sweeps(n, flow):
        if flow:
            os.environ.pop("CT2_ART_FLOW", None)
        else:
            os.environ["CT2_ART_FLOW"] = "0"
        x = torch.zeros(vol.shape, dtype=torch.float32, device="cuda")
        x = vp.artstep_linear(0.5, sino, None, fvd, *geo, x, *g2)   # warm-up
        x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            x = vp.artstep_linear(0.5, sino, None, fvd, *geo, x, *g2)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        os.environ.pop("CT2_ART_FLOW", None)
        return t, x.cpu().numpy()

    t_flow, x_flow = sweeps(10, True)
    t_level, x_level = sweeps(10, False)
    line = {"metric": "seconds per 10 ART sweeps (64^3 Shepp-Logan, 90 views, 96^2 detector, "
                      "relaxation 0.5)",
            "value": round(t_flow, 4), "unit": "s", "n_gpus": 1, "steps": 10, "warmup": 1,
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic: 10-ellipsoid Shepp-Logan phantom and its "
                                    "projections",
            "config": {"workload": "C1: ART from zero on the phantom's cone-beam projections"},
            "ms_per_sweep": {"dataflow": round(1e3 * t_flow / 10, 2),
                             "level_schedule": round(1e3 * t_level / 10, 2)},
            "dataflow_equals_level_schedule": bool(np.array_equal(x_flow, x_level))}
    if not args.no_cpu:
        mods, kind = ((oracle.reference(), "reference") if oracle.ref_available()
                      else (oracle.restated("f32"), "port"))
        x0 = np.zeros_like(vol)
        t0 = time.perf_counter()
        x1 = mods.view_proj3D.artstep_linear(0.5, sino_h, None, fv, *geo, x0, *g2)
        t1 = time.perf_counter() - t0
        _, g1 = sweeps(1, True)
        line["cpu_baseline"] = {"value": round(10 * t1, 3), "unit": "s", "cores": 1,
                                "kind": kind, "sample": f"one sweep ({t1:.2f} s) x 10, single "
                                "thread" + (", reference Fortran compiled by amdflang -O2 from "
                                            "emil_code/view_proj3D.f" if kind == "reference"
                                            else ", FP32 C restatement")}
        line["gpu_over_cpu"] = round(10 * t1 / t_flow, 1)
        line["rel_l2_one_sweep_vs_cpu"] = float(np.linalg.norm(g1 - x1) / np.linalg.norm(x1))
    print(json.dumps(line), flush=True)


def asd_workload(args):
    """ASD-POCS at 256^3: the recon workload's noisy data; 10 outer
    iterations of a 10-subset SART sweep, the two distances, TV value, the
    data distance (one forward) and 20 normalised TV-gradient steps, with gf,
    reg_fac and meps adapted on the device; the first iteration eager, the
    other 9 replays of one captured HIP graph."""
    import numpy as np
    import torch
    from ct2 import tomo3D
    from ct2.operator import ConeBeamOperator
    from ct2.solvers import AsdPocs
    mu = 0.02
    hw = N * VOX / 2
    sino = tomo3D.sinogram3D("circular_conebeam", {"radius": R, "source_to_detector": SDD},
                             (NS, ND, ND), 2 * np.pi, ND * PITCH, ND * PITCH, 0.0,
                             -ND * PITCH / 2, -ND * PITCH / 2, 0)
    img = tomo3D.image3D((N, N, N), 2 * hw, 2 * hw, 2 * hw, -hw, -hw, -hw)
    op = ConeBeamOperator.from_objects(sino, img, arith="fast")
    truth = torch.from_numpy(phantom()).cuda() * mu
    p = op.forward(truth)
    g = torch.Generator(device="cuda").manual_seed(20260102)
    counts = torch.poisson(1e5 * torch.exp(-p.double()), generator=g)
    b = (-torch.log(torch.clamp(counts, min=1.0) / 1e5)).float().contiguous()
    x = torch.zeros((N, N, N), dtype=torch.float32, device="cuda")
    t0 = time.perf_counter()
    st = AsdPocs(op, b, x, nsub=10, ngrad=20, truth=truth, max_iter=20)
    torch.cuda.synchronize()
    t_plan = time.perf_counter() - t0
    t0 = time.perf_counter()
    rec1, _ = st.run(1, graph=False)
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    rec, _ = st.run(9, graph=True, save_every=9)
    torch.cuda.synchronize()
    t_loop = time.perf_counter() - t0
    rec = np.concatenate([rec1, rec], axis=0)
    tn = float(torch.linalg.vector_norm(truth))
    upd_iter = 3 * op.updates_per_direction()     # SART forward + adjoint, data-distance forward
    line = {"metric": "ms per ASD-POCS iteration (10-subset SART + 20 TV steps, 256^3, 360 "
                      "views, 512^2 detector)",
            "value": round(1e3 * t_loop / 9, 2), "unit": "ms", "n_gpus": 1, "steps": 9,
            "warmup": 1, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32",
            "data": "synthetic: 40-ellipsoid phantom x 0.02/mm, Poisson noise 1e5 photons/ray",
            "config": {"workload": "ASD-POCS, SART 10 subsets, 20 TV steps, graph replays"},
            "phases_s": {"plan": round(t_plan, 3), "first_eager": round(t_first, 3),
                         "9_graph_replays": round(t_loop, 3)},
            "gupd_per_s": round(9 * upd_iter / t_loop / 1e9, 1),
            "records": {"data_distance": [round(float(v), 4) for v in rec[:, 1]],
                        "rel_l2_to_truth": [round(float(v) / tn, 4) for v in rec[:, 3]]}}
    print(json.dumps(line), flush=True)


def _time_ms(fn, n=5):
    """Best of n calls of fn, each bracketed by HIP events on the current stream."""
    import torch
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(n):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
    return min(ts)


def tomo_workload(args):
    """The limited-angle tomosynthesis workload: 1024x768x64 volume (0.1 x 0.1
    x 1 mm) of the breast-like phantom (seed 20260103), 25 views over +-25 deg
    (R = 650 mm, detector 20 mm below the centre of rotation), 2048x1536
    detector at 0.1 mm, z-driven projector and its transpose.  Times the
    forward with f32, f16 and bf16 volume storage, C-ordered and plane-major,
    and the adjoint into f32, in fast and in reference arithmetic (best of 5
    calls each); the value is the f16 plane-major forward plus the fast
    adjoint; the
    precision of each output against FP64 is tests/test_gpu_tomo_precision's
    table."""
    import numpy as np
    import torch
    from ct2 import configs3D, phantoms
    from ct2.operator import ConeBeamOperator, to_storage
    nx, ny, nz, vox = 1024, 768, 64, (0.1, 0.1, 1.0)
    ns, nu, nv, pitch = 25, 2048, 1536, 0.1
    log("building the tomosynthesis phantom")
    fv = configs3D.frame_vectors("tomosynthesis", {"radius": 650.0, "center_to_detector": 20.0},
                                 ns, np.deg2rad(50.0), np.deg2rad(65.0), 1)
    vol = torch.from_numpy(phantoms.breast_like((nx, ny, nz), vox, 20260103)).cuda()
    ops = {a: ConeBeamOperator(fv, ns, nu, nv, pitch, pitch, -nu * pitch / 2, -nv * pitch / 2,
                               (nx, ny, nz), vox, (-nx * vox[0] / 2, -ny * vox[1] / 2, 0.0),
                               axis="z", arith=a) for a in ("fast", "exact")}
    sino = torch.empty(ops["fast"].sino_shape, dtype=torch.float32, device="cuda")
    out = torch.empty((nx, ny, nz), dtype=torch.float32, device="cuda")
    v16 = to_storage(vol, torch.float16, "stochastic")
    vb16 = to_storage(vol, torch.bfloat16, "stochastic")
    # plane-major copies (nz planes of C-ordered (x, y)): a wavefront's rays
    # then gather neighbouring (x, y) of one plane from contiguous memory
    z32 = to_storage(vol, torch.float32, zplanes=True)
    z16 = to_storage(vol, torch.float16, "stochastic", zplanes=True)
    ops["fast"].forward(vol, out=sino)
    y = sino.clone()
    upd = ns * nu * nv * nz   # z-driven: one sample per ray and z plane
    ms = {
        "fwd_fast_f32": _time_ms(lambda: ops["fast"].forward(vol, out=sino)),
        "fwd_fast_f16": _time_ms(lambda: ops["fast"].forward(v16, out=sino)),
        "fwd_fast_bf16": _time_ms(lambda: ops["fast"].forward(vb16, out=sino)),
        "fwd_exact_f32": _time_ms(lambda: ops["exact"].forward(vol, out=sino)),
        "fwd_fast_f32_zplanes": _time_ms(lambda: ops["fast"].forward(z32, out=sino, layout="zxy")),
        "fwd_fast_f16_zplanes": _time_ms(lambda: ops["fast"].forward(z16, out=sino, layout="zxy")),
        "fwd_exact_f32_zplanes": _time_ms(
            lambda: ops["exact"].forward(z32, out=sino, layout="zxy")),
        "bp_fast_f32": _time_ms(lambda: ops["fast"].adjoint(y, out=out)),
        "bp_exact_f32": _time_ms(lambda: ops["exact"].adjoint(y, out=out)),
    }
    gupd = {k: round(upd / (v * 1e-3) / 1e9, 1) for k, v in ms.items()}
    mixed = ms["fwd_fast_f16_zplanes"] + ms["bp_fast_f32"]
    fp32 = ms["fwd_exact_f32"] + ms["bp_exact_f32"]
    line = {"metric": "Gvox-ray updates/s (tomosynthesis fwd+bp, 1024x768x64, 25 views, "
                      "2048x1536 detector), f16 storage vs reference arithmetic",
            "value": round(2 * upd / (mixed * 1e-3) / 1e9, 1), "unit": "Gvox-ray updates/s",
            "n_gpus": 1, "steps": 5, "warmup": 1, "ms_per_step": round(mixed, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "storage": "f16 plane-major volume (forward), f32 sinogram and adjoint",
            "data": "synthetic: breast-like phantom (200 ellipsoids, 50 calcifications), "
                    "its own projections as the adjoint's input",
            "config": {"workload": "C3: limited-angle tomosynthesis, z-driven projector"},
            "kernels_ms": {k: round(v, 3) for k, v in ms.items()}, "gupd_per_s": gupd,
            "fp32_value": round(2 * upd / (fp32 * 1e-3) / 1e9, 1),
            "mixed_over_fp32": round(fp32 / mixed, 2)}
    print(json.dumps(line), flush=True)


def adam_workload(args):
    """The differentiation workload: 128^3 random ellipsoids (seed 20260105),
    120 views over 360 deg, 192^2 detector at 1.6 mm (R = 500, SDD = 1000);
    1/2||Ax-b||^2 + 1e-3 TV(x) minimised by torch Adam (lr 1e-2) for 200
    steps through ct2.autograd (fast arithmetic); one step is a forward, the
    loss, its backward (adjoint + TV gradient) and the Adam update."""
    import numpy as np
    import torch
    from ct2 import configs3D, phantoms
    from ct2.operator import ConeBeamOperator
    from ct2.solvers import tv_least_squares
    n, ns, nd = 128, 120, 192
    fv = configs3D.frame_vectors("circular_conebeam", {"radius": 500.0, "source_to_detector": 1000.0},
                                 ns, 2 * np.pi, 0.0, 0)
    op = ConeBeamOperator(fv, ns, nd, nd, 1.6, 1.6, -nd * 0.8, -nd * 0.8, (n, n, n),
                          (1.0, 1.0, 1.0), (-n / 2, -n / 2, -n / 2), arith="fast")
    truth = torch.from_numpy(np.ascontiguousarray(phantoms.random_ellipsoids(n, n / 2, 40, 20260105),
                                                  np.float32)).cuda()
    b = op.forward(truth)
    tv_least_squares(op, b, lam=1e-3, steps=5, lr=1e-2)   # warm-up (tables, workspaces)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, losses = tv_least_squares(op, b, lam=1e-3, steps=200, lr=1e-2)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    tn = float(torch.linalg.vector_norm(truth))
    err = float(torch.linalg.vector_norm(x - truth)) / tn
    upd = 2 * op.updates_per_direction()
    line = {"metric": "ms per Adam step (1/2||Ax-b||^2 + 1e-3 TV, 128^3, 120 views, 192^2 "
                      "detector, through ct2.autograd)",
            "value": round(1e3 * t / 200, 3), "unit": "ms", "n_gpus": 1, "steps": 200,
            "warmup": 5, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic: 40-ellipsoid phantom and its projections",
            "config": {"workload": "C5: TV-regularised least squares by torch Adam"},
            "gupd_per_s": round(200 * upd / t / 1e9, 1),
            "loss_first_last": [losses[0], losses[-1]], "rel_l2_to_truth": round(err, 4)}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU leg (baseline timing and FP64 check)")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 precision check")
    ap.add_argument("--profile-once", action="store_true",
                    help="for counter runs: one warm-up and one mixed fwd+bp step, then a "
                         "512 MiB device copy that calibrates FETCH_SIZE/WRITE_SIZE; no JSON")
    ap.add_argument("--workload", choices=["c2", "batch", "recon", "asd", "art", "tomo", "adam"],
                    default="c2",
                    help="c2: the headline fwd+bp metric (default); batch: 64-problem batched "
                         "SART in a HIP graph; recon: FDK + 20 EM-TV iterations at 256^3; "
                         "asd: ASD-POCS iterations (SART + TV) in a HIP graph at 256^3; "
                         "art: C1, 10 ART sweeps on the GPU beside the reference Fortran; "
                         "tomo: C3, tomosynthesis fwd and bp per storage type; adam: C5, "
                         "200 Adam steps through the autograd projector")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the "
              "multi-GPU sharding is a design only: views for A, a z-slab reduce-scatter "
              "for A^T)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)
    if args.workload == "batch":
        batch_workload(args)
        return
    if args.workload == "asd":
        asd_workload(args)
        return
    if args.workload == "recon":
        recon_workload(args)
        return
    if args.workload == "art":
        art_workload(args)
        return
    if args.workload == "tomo":
        tomo_workload(args)
        return
    if args.workload == "adam":
        adam_workload(args)
        return
    fv, geo, grid = geometry()
    log("building phantom")
    vol_np = phantom()
    vol32 = torch.from_numpy(vol_np).cuda()
    # f16 volume storage: the column-separable forward reads whole z rows of
    # it (2 B per voxel)
    vol16 = torch.empty(tuple(vol32.shape), dtype=torch.float16, device="cuda")
    op_mixed = make_op(fv, geo, grid, "fast")
    op_fp32 = make_op(fv, geo, grid, "exact")
    sino = torch.empty(op_mixed.sino_shape, dtype=torch.float32, device="cuda")
    recon = torch.empty(op_mixed.shape, dtype=torch.float32, device="cuda")
    nupd = 2 * op_mixed.updates_per_direction()
    data = noisy_data(op_fp32, vol32)

    if args.profile_once:
        for _ in range(2):
            to_storage(vol32, vol16, stochastic=True)
            op_mixed.forward(vol16, out=sino)
            op_mixed.adjoint(data, out=recon)
        src = torch.zeros(128 * 1024 * 1024, dtype=torch.float32, device="cuda")   # 512 MiB
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("profile step done")
        return

    log("timing mixed path (FP16 storage with stochastic rounding, fast arithmetic)")
    wall_m, tc_m, tf_m, tb_m = time_path(op_mixed, vol32, vol16, sino, data, recon, args.steps,
                                         args.warmup)
    log(f"  convert {tc_m:.3f} ms  fwd {tf_m:.2f} ms  bp {tb_m:.2f} ms  "
        f"wall/step {1e3 * wall_m / args.steps:.2f} ms")
    log("timing FP32-fast path (FP32 storage, fast arithmetic)")
    sino32 = torch.empty_like(sino)
    wall_q, _, tf_q, tb_q = time_path(op_mixed, vol32, None, sino32, data, recon,
                                      max(1, args.steps // 2), 1)
    log(f"  fwd {tf_q:.2f} ms  bp {tb_q:.2f} ms")
    log("timing FP32 path (FP32 storage, reference operation order)")
    wall_f, _, tf_f, tb_f = time_path(op_fp32, vol32, None, sino32, data, recon,
                                      max(1, args.steps // 2), 1)
    log(f"  fwd {tf_f:.2f} ms  bp {tb_f:.2f} ms")
    sub = check_views(NS)
    gpu_out = None if args.no_check else gpu_precision_outputs(op_mixed, op_fp32, vol32, vol16,
                                                               sub)

    value = nupd / (wall_m / args.steps) / 1e9
    fast32_value = nupd / (wall_q / max(1, args.steps // 2)) / 1e9
    fp32_value = nupd / (wall_f / max(1, args.steps // 2)) / 1e9
    # the dominant part of the mixed step is the adjoint call (the
    # column-separable route: every view of the scan has e_v along z); its
    # kernels' counts are summed in the profile, and its time is the call's
    kname = ("ct2_backproject, column-separable route: k_sep_tables + k_bp_sep<4> + "
             "k_sep_zero + brick pass (f32 out)")
    t_k = tb_m
    cnt = read_counters()
    rates = issue_rates()
    log("VALU issue rates (wavefront instructions/s): " +
        ", ".join(f"{k} {v:.3e}" for k, v in rates.items()))
    valu, v64 = cnt.get("valu_insts_per_launch"), cnt.get("valu_f64_insts_per_launch")
    roof = {"bound": "valu", "kernel": kname, "kernel_ms": round(t_k, 3),
            "unit": "Tlane-op/s", "peak": None, "achieved": None, "frac": None,
            "valu_insts_per_launch": valu, "valu_f64_insts_per_launch": v64,
            "issue_rates_per_cu_clock": {k: round(v / (256 * 2.4e9), 3)
                                         for k, v in rates.items()},
            "published_fp32_vector_tflops": 157.3}
    if valu and v64 is not None:
        # the kernel's instruction mix at the measured issue rates: its FP64
        # instructions at the FP64 rate, the others at the FP32 rate
        t_min = v64 / rates["f64_fma"] + (valu - v64) / rates["f32_fma"]
        ach = valu * 64 / (t_k * 1e-3) / 1e12
        peak = valu * 64 / t_min / 1e12
        roof.update({"achieved": round(ach, 3), "peak": round(peak, 3),
                     "frac": round(ach / peak, 4), "valu_bound_ms": round(1e3 * t_min, 3)})
    # compulsory HBM bytes of the adjoint: the sinogram and indicator read
    # once, the f32 volume written once
    ns_, nu_, nv_ = op_mixed.sino_shape
    comp = ns_ * nu_ * nv_ * 5 + recon.numel() * 4
    traffic = cnt.get("traffic_bytes_per_launch")
    roof["hbm"] = {"peak": HBM_PEAK_GBS, "unit": "GB/s", "compulsory_bytes": comp,
                   "compulsory_GBs": round(comp / (t_k * 1e-3) / 1e9, 1),
                   "traffic": traffic,
                   "measured_GBs": round(traffic / (t_k * 1e-3) / 1e9, 1) if traffic else None,
                   "frac": round(traffic / (t_k * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                   if traffic else None}
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gvox-ray updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall_m / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "storage": "f16 volume (forward, 2 B per voxel), f32 sinogram, f32 volume out "
                   "(adjoint)",
        "data": "synthetic: seeded 40-ellipsoid phantom, views of a circular cone-beam scan; "
                "the adjoint is applied to its Poisson-noised projections (1e5 photons/ray, "
                "seed 20260102), where no ray is zero",
        "config": {"workload": "Joseph forward + exact-adjoint backprojection, 256^3 volume "
                               "(0.8 mm), 360 views, 512x512 detector (0.8 mm), R=500 mm, "
                               "SDD=1000 mm",
                   "updates_per_step": nupd},
        "kernels_ms": {"convert_mixed": round(tc_m, 4), "fwd_mixed": round(tf_m, 3),
                       "bp_mixed": round(tb_m, 3),
                       "fwd_fast_f32": round(tf_q, 3), "bp_fast_f32": round(tb_q, 3),
                       "fwd_fp32": round(tf_f, 3), "bp_fp32": round(tb_f, 3)},
        "fast_f32_value": round(fast32_value, 3),
        "fp32_value": round(fp32_value, 3),
        "mixed_over_fast_f32": round(value / fast32_value, 3),
        "mixed_over_fp32": round(value / fp32_value, 3),
        "rel_l2_vs_fp64": None,
        "roofline": roof,
    }
    if not args.no_cpu:
        log("CPU leg: reference path on a bounded sample; FP64 oracle check of the GPU outputs")
        base, prec = cpu_leg(fv, geo, grid, vol_np, gpu_out=gpu_out, sub=sub)
        line["cpu_baseline"] = base
        line["cpu_baseline_multicore"] = cpu_multicore(fv, geo, grid, vol_np)
        line["rel_l2_vs_fp64"] = prec
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
