"""Iteration drivers built on the device operators.

* ``em_tv`` -- MLEM alternated with normalised TV-gradient descent, the
  ASD-POCS pattern of emil_code/examples/jawPhantomCCB.py:191-327 with the
  EM step (tomo3D.py:1394-1435) in place of ART: per outer iteration one EM
  update, positivity, then ``n_tv`` steps x <- max(x - gf*g/|g|, 0) with
  g = total_var_grad3d(x) (totalvar.f:476-602).  gf starts as
  ``alpha * |x_EM - x_prev|`` (jawPhantomCCB.py:170, :213-215) and is reduced
  by ``red_fact`` whenever the TV descent moved the image at least ``conv``
  times as far as the data step (:321-322).  Every quantity stays on the
  device (distances by ct2_distance, step sizes as 0-d tensors), so one
  outer iteration can be captured into a HIP graph and replayed.
* ``asd_pocs`` -- the jawPhantomCCB.py driver itself on tomo3D objects
  (ART sweeps, zero_edges, make_positive, TV descent, adaptive gf and
  reg_fac, data-error tolerance at ``iter_eps``), with a checkpoint that
  resumes correctly (the reference reads gf, reg_fac and meps from one
  ``read()`` of parms.txt, jawPhantomCCB.py:160-162, so its resume fails,
  and saves in the middle of an iteration) and the traffic-file stop of
  :235-244.
* ``tv_least_squares`` -- 1/2||Ax-b||^2 + lam*TV(x) minimised by torch.optim
  through the autograd projector (ct2.autograd).
"""
from __future__ import annotations

import os

import numpy as np
import torch

from . import _dev
from ._lib import CT2_TV_WORK_DOUBLES, call
from .autograd import project, total_variation, tv_grad_dev, tv_value_dev


def distance(a, b=None, mask=None, out=None, work=None):
    """sqrt(sum mask*(a-b)^2) in FP64 on the device, as a (1,) float64 tensor."""
    out = torch.empty((1,), dtype=torch.float64, device=a.device) if out is None else out
    work = torch.empty(CT2_TV_WORK_DOUBLES, dtype=torch.float64, device=a.device) \
        if work is None else work
    call("ct2_distance", _dev.ptr(a), _dev.ptr(b), _dev.ptr(mask), a.numel(), _dev.ptr(out),
         _dev.ptr(work), _dev.stream_handle())
    return out


class EmTv:
    """State and one outer iteration of ``em_tv`` with every buffer
    preallocated, so that ``iterate`` launches no allocation of its own and
    can be captured into a torch.cuda.CUDAGraph."""

    def __init__(self, op, sino, x, n_tv=5, eps=1e-8, alpha=0.2, red_fact=0.98, conv=0.99):
        self.op, self.sino, self.x = op, sino, x
        self.n_tv, self.eps, self.alpha = int(n_tv), float(eps), float(alpha)
        self.red_fact, self.conv = float(red_fact), float(conv)
        dev = x.device
        n_work = 2 * sino.numel() + 2 * x.numel()
        self.work = torch.empty(n_work, dtype=torch.float32, device=dev)
        self.sens = op.sensitivity()
        self.prev = torch.empty_like(x)
        self.g = torch.empty_like(x)
        self.bobo = torch.empty((1,), dtype=torch.float32, device=dev)
        self.rwork = torch.empty(CT2_TV_WORK_DOUBLES, dtype=torch.float64, device=dev)
        self.adist = torch.zeros((1,), dtype=torch.float64, device=dev)
        self.gdist = torch.zeros((1,), dtype=torch.float64, device=dev)
        self.gf = torch.zeros((1,), dtype=torch.float32, device=dev)
        self.started = torch.zeros((1,), dtype=torch.float32, device=dev)

    def iterate(self):
        x = self.x
        self.prev.copy_(x)
        self.op.em_step(self.sino, x, work=self.work)
        x.clamp_(min=0.0)
        distance(x, self.prev, out=self.adist, work=self.rwork)
        # first iteration: gf <- alpha*adist; afterwards gf is kept
        a32 = self.adist.to(torch.float32)
        self.gf.copy_(torch.where(self.started > 0, self.gf, self.alpha * a32))
        self.started.fill_(1.0)
        self.prev.copy_(x)
        for _ in range(self.n_tv):
            tv_grad_dev(x, 0.0, self.eps, normalizeq=1, out=self.g, bobo=self.bobo,
                        work=self.rwork)
            x.sub_(self.gf * self.g)
            x.clamp_(min=0.0)
        distance(x, self.prev, out=self.gdist, work=self.rwork)
        shrink = self.gdist >= self.conv * self.adist
        self.gf.mul_(torch.where(shrink, self.red_fact, 1.0).to(torch.float32))


def em_tv(op, sino, x, n_em=20, n_tv=5, eps=1e-8, alpha=0.2, graph=False):
    """Run ``n_em`` outer iterations in place on x; with ``graph=True`` the
    first iteration runs eagerly (it also tunes the adjoint) and the rest are
    replays of one captured iteration."""
    st = EmTv(op, sino, x, n_tv=n_tv, eps=eps, alpha=alpha)
    if n_em <= 0:
        return x
    st.iterate()
    if not graph:
        for _ in range(n_em - 1):
            st.iterate()
        return x
    g = capture(st.iterate)
    for _ in range(n_em - 1):
        g.replay()
    return x


def capture(fn):
    """Capture fn (device work only, no host synchronisation, already run
    once eagerly so that lazy set-up and the adjoint's brick tuning are done)
    into a graph; returns the torch.cuda.CUDAGraph to replay."""
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        fn()
    return g


def fdk_em_tv(sino, image, op, n_em=20, n_tv=5, eps=1e-8, graph=True):
    """FDK reconstruction followed by ``em_tv`` from max(FDK, 0) on device
    tensors.  ``sino`` is a tomo3D.sinogram3D (circular cone beam), ``image``
    the tomo3D.image3D geometry; returns the (nx, ny, nz) device volume."""
    from . import FDK
    b = _dev.to_dev(sino.mat, torch.float32, (sino.ns, sino.nu, sino.nv))
    x = FDK.fdk_reconstruct(b, sino, image).contiguous()
    x.clamp_(min=0.0)
    em_tv(op, b, x, n_em=n_em, n_tv=n_tv, eps=eps, graph=graph)
    return x


# --------------------------------------------------------------- ASD-POCS --
def _save_checkpoint(path, it, gf, reg_fac, meps, image):
    """The reference's run state: imageTemp.dat, iternum.txt, parms.txt
    (examples/sphericalRadon.py:171-183)."""
    from . import Utils
    Utils.write_state(path, it, gf, reg_fac, meps, image.mat)


def _load_checkpoint(path, image):
    from . import Utils
    it, gf, reg_fac, meps = Utils.read_state(path)
    arr = Utils.read_dat(os.path.join(path, "imageTemp.dat"), (image.nx, image.ny, image.nz))
    if isinstance(image.mat, torch.Tensor):
        image.mat = torch.from_numpy(arr).to(image.mat.device)
    else:
        image.mat = arr
    return it, gf, reg_fac, meps


def asd_pocs(image, sino, work_sino, max_iter, nart=1, ngradmax=20, gf=0.2, reg_fac=1.0,
             red_fact=0.98, red_reg=0.995, conv=0.99, iter_eps=500, art_alg='art_linear',
             checkpoint_dir=None, resume=False, traffic_file=None, truth=None, log=None,
             bang_num=None):
    """ASD-POCS of emil_code/examples/jawPhantomCCB.py:191-327 on tomo3D
    objects (numpy or device arrays).  Returns the list of per-iteration
    records (iter, data distance, TV, image error).  With checkpoint_dir the
    records are also appended to derrEvol.txt there, the run state is saved
    after every iteration in the reference's files (imageTemp.dat,
    iternum.txt, parms.txt), and with bang_num a snapshot image<iter>.dat is
    written every bang_num iterations (jawPhantomCCB.py:257-268)."""
    meps = -1.0
    it = 0
    if resume:
        it, gf, reg_fac, meps = _load_checkpoint(checkpoint_dir, image)
    pre = image.duplicate()
    gimage = image.duplicate()
    records = []
    done = False
    while not done and it < max_iter:
        it += 1
        image.copy_to(pre)
        for _ in range(nart):
            image.art_step(sino, reg_fac, alg=art_alg)
        image.zero_edges()
        image.make_positive()
        adist = image.dist_to(pre)
        if it == 1:
            gf *= adist
        if traffic_file is not None and os.path.exists(traffic_file):
            with open(traffic_file) as f:
                done = f.read().strip() == "stop"
        posttv = image.calc_total_var(0.0)
        image.project_to(work_sino)
        ddist = sino.dist_to(work_sino)
        if it == iter_eps:
            meps = ddist
        ierr = image.dist_to(truth) if truth is not None else float("nan")
        if not done:
            records.append((it, ddist, posttv, ierr))
            if log is not None:
                log(it, ddist, posttv, ierr)
            if checkpoint_dir is not None:
                from . import Utils
                Utils.append_derr(checkpoint_dir, it, ddist, posttv, ierr)
        if checkpoint_dir is not None and bang_num and it % bang_num == 0:
            from . import Utils
            Utils.write_dat(os.path.join(checkpoint_dir, f"image{it}.dat"), image.mat)
        image.copy_to(pre)
        ngrad = 0
        while ngrad < ngradmax and not done:
            ngrad += 1
            gimage.mat = image.mat
            gimage.tv_grad()
            gimage.scale(gf)
            image.mat = image.mat - gimage.mat
            image.zero_edges()
            image.make_positive()
        gdist = image.dist_to(pre)
        if gdist >= conv * adist and ddist > meps:
            gf *= red_fact
        reg_fac *= red_reg
        # saved after the whole iteration, so a resumed run continues the
        # same trajectory (the reference saves before its TV descent)
        if checkpoint_dir is not None:
            _save_checkpoint(checkpoint_dir, it, gf, reg_fac, meps, image)
    return records


# ------------------------------------------------ ASD-POCS in a HIP graph --
def _zero_edges_(x):
    """Two-voxel border of a device volume set to zero (tomo3D.py:928-943)."""
    for sl in ((slice(0, 2),), (slice(-2, None),), (slice(None), slice(0, 2)),
               (slice(None), slice(-2, None)), (slice(None), slice(None), slice(0, 2)),
               (slice(None), slice(None), slice(-2, None))):
        x[sl].zero_()


class AsdPocs:
    """ASD-POCS (examples/jawPhantomCCB.py:191-327) on device tensors with the
    data step done by ordered-subset SART over a one-problem BatchPlan, and
    every scalar of the adaptive schedule (gf, reg_fac, meps, the distances,
    the iteration counter) kept on the device, so that one outer iteration
    contains no host synchronisation and is captured into a HIP graph.

    One iteration: SART sweep over nsub subsets with relaxation reg_fac;
    zero_edges, positivity; adist = |x - x_pre|; gf *= adist on the first
    iteration; data distance |A x - b| over used rays and TV (the record the
    reference writes to derrEvol.txt); ngrad normalised TV-gradient steps of
    size gf, each followed by zero_edges and positivity; gdist = |x - x_pre|;
    gf *= red_fact when gdist >= conv*adist and the data distance is above
    meps (the distance at iteration iter_eps); reg_fac *= red_reg.  The
    records accumulate in a device log; ``run`` writes them and the run
    state (the reference's files) between replays and checks the traffic
    file there."""

    def __init__(self, op, sino, x, nsub=10, ngrad=20, gf=0.2, reg_fac=1.0, red_fact=0.98,
                 red_reg=0.995, conv=0.99, iter_eps=500, max_iter=1000, truth=None, eps=1e-9):
        from .batch import BatchPlan
        if x.dtype != torch.float32 or not x.is_contiguous() or not x.is_cuda:
            raise ValueError("x must be a contiguous float32 device tensor")
        self.op, self.b, self.x, self.truth = op, sino, x, truth
        self.ngrad, self.eps = int(ngrad), float(eps)
        self.red_fact, self.red_reg, self.conv = float(red_fact), float(red_reg), float(conv)
        self.iter_eps = int(iter_eps)
        dev = x.device
        self.plan = BatchPlan([op], [sino], [x], nsub=nsub)
        self.pre = torch.empty_like(x)
        self.g = torch.empty_like(x)
        self.proj = torch.empty(op.sino_shape, dtype=torch.float32, device=dev)
        self.bobo = torch.empty((1,), dtype=torch.float32, device=dev)
        self.rwork = torch.empty(CT2_TV_WORK_DOUBLES, dtype=torch.float64, device=dev)
        f64 = dict(dtype=torch.float64, device=dev)
        self.adist, self.gdist = torch.zeros((1,), **f64), torch.zeros((1,), **f64)
        self.ddist, self.tv = torch.zeros((1,), **f64), torch.zeros((1,), **f64)
        self.ierr = torch.full((1,), float("nan"), **f64)
        self.meps = torch.full((1,), -1.0, **f64)
        self.gf = torch.full((1,), float(gf), dtype=torch.float32, device=dev)
        self.reg_fac = torch.full((1,), float(reg_fac), dtype=torch.float32, device=dev)
        self.started = torch.zeros((1,), dtype=torch.float32, device=dev)
        self.it = torch.zeros((1,), dtype=torch.int64, device=dev)
        self.log = torch.zeros((int(max_iter) + 1, 4), **f64)
        self.row = torch.zeros((1, 4), **f64)
        self._graph = None

    def iterate(self):
        x = self.x
        self.it.add_(1)
        self.pre.copy_(x)
        call("ct2_batch_sart_dev", self.plan._h, _dev.ptr(self.reg_fac), 1, _dev.stream_handle())
        _zero_edges_(x)
        x.clamp_(min=0.0)
        distance(x, self.pre, out=self.adist, work=self.rwork)
        self.gf.copy_(torch.where(self.started > 0, self.gf,
                                  self.gf * self.adist.to(torch.float32)))
        self.started.fill_(1.0)
        tv_value_dev(x, 0.0, self.eps, out=self.tv, work=self.rwork)
        self.op.forward(x, out=self.proj)
        distance(self.proj, self.b, mask=self.op.ind, out=self.ddist, work=self.rwork)
        self.meps.copy_(torch.where(self.it == self.iter_eps, self.ddist, self.meps))
        if self.truth is not None:
            distance(x, self.truth, out=self.ierr, work=self.rwork)
        self.row[:, 0].copy_(self.it.to(torch.float64))
        self.row[:, 1].copy_(self.ddist)
        self.row[:, 2].copy_(self.tv)
        self.row[:, 3].copy_(self.ierr)
        self.log.index_copy_(0, self.it.clamp(max=self.log.shape[0] - 1), self.row)
        self.pre.copy_(x)
        for _ in range(self.ngrad):
            tv_grad_dev(x, 0.0, self.eps, normalizeq=1, out=self.g, bobo=self.bobo,
                        work=self.rwork)
            x.sub_(self.gf * self.g)
            _zero_edges_(x)
            x.clamp_(min=0.0)
        distance(x, self.pre, out=self.gdist, work=self.rwork)
        shrink = (self.gdist >= self.conv * self.adist) & (self.ddist > self.meps)
        self.gf.mul_(torch.where(shrink, self.red_fact, 1.0).to(torch.float32))
        self.reg_fac.mul_(self.red_reg)

    def state(self):
        """(iteration, gf, reg_fac, meps) as host numbers (synchronises)."""
        return (int(self.it.item()), float(self.gf.item()), float(self.reg_fac.item()),
                float(self.meps.item()))

    def set_state(self, it, gf, reg_fac, meps):
        self.it.fill_(int(it))
        self.gf.fill_(float(gf))
        self.reg_fac.fill_(float(reg_fac))
        self.meps.fill_(float(meps))
        self.started.fill_(1.0 if it > 0 else 0.0)

    def records(self, first, last):
        """Log rows of iterations first..last (inclusive) as a host array."""
        return self.log[first:last + 1].cpu().numpy()

    def run(self, n_iter, graph=True, checkpoint_dir=None, save_every=10, traffic_file=None):
        """n_iter more iterations (replays of one captured iteration after an
        eager first one).  Every save_every iterations, and at the end, the
        new log rows are appended to derrEvol.txt and the run state is saved
        in checkpoint_dir; the traffic file (``stop``) is read there too."""
        from . import Utils
        it0 = self.state()[0]
        done_it = it0
        stopped = False
        for k in range(n_iter):
            if k == 0 or not graph:
                self.iterate()
                if graph and self._graph is None and n_iter > 1:
                    self._graph = capture(self.iterate)
                    # capture runs nothing; the captured iteration is replayed below
            else:
                self._graph.replay()
            last = k == n_iter - 1
            if (k + 1) % max(int(save_every), 1) == 0 or last:
                torch.cuda.synchronize()
                it = self.state()[0]
                if checkpoint_dir is not None:
                    for r in self.records(done_it + 1, it):
                        Utils.append_derr(checkpoint_dir, int(r[0]), r[1], r[2], r[3])
                    Utils.write_state(checkpoint_dir, *self.state(), self.x)
                done_it = it
                if traffic_file is not None and os.path.exists(traffic_file):
                    with open(traffic_file) as f:
                        if f.read().strip() == "stop":
                            stopped = True
                            break
        return self.records(it0 + 1, self.state()[0]), stopped

    def resume(self, checkpoint_dir):
        """Loads the reference's run-state files into x and the schedule."""
        from . import Utils
        it, gf, reg_fac, meps = Utils.read_state(checkpoint_dir)
        arr = Utils.read_dat(os.path.join(checkpoint_dir, "imageTemp.dat"), tuple(self.x.shape))
        self.x.copy_(torch.from_numpy(arr).to(self.x.device))
        self.set_state(it, gf, reg_fac, meps)


# ------------------------------------------------ TV-regularised LS (Adam) --
def tv_least_squares(op, b, lam=1e-3, steps=200, lr=1e-2, x0=None, eps=1e-8, callback=None):
    """min_x 1/2||Ax-b||^2 + lam*TV(x) by torch Adam through the exact
    projector/adjoint pair; returns (x, list of losses)."""
    x = torch.zeros(op.shape, dtype=torch.float32, device=b.device) if x0 is None \
        else x0.detach().clone()
    x.requires_grad_(True)
    opt = torch.optim.Adam([x], lr=lr)
    losses = []
    for k in range(steps):
        opt.zero_grad(set_to_none=False)
        r = project(x, op) - b
        loss = 0.5 * (r * r).sum(dtype=torch.float64).to(torch.float32) + \
            lam * total_variation(x, 0.0, eps)
        loss.backward()
        opt.step()
        losses.append(loss.detach())
        if callback is not None:
            callback(k, x, loss)
    return x.detach(), [float(v) for v in losses]
