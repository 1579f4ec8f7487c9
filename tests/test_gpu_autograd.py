"""Differentiation through the HIP projector (ct2.autograd).

* dot-product test <Ax, y> = <x, A^T y> through autograd on Gaussian x, y;
* the autograd gradient of 1/2||Ax-b||^2 against the reference Fortran's
  backprojection (golden fixture written from emil_code/view_proj3D.f);
* central finite differences in FP64 on 64 random voxels (step 1e-3) of
  1/2||Ax-b||^2 + lam*TV(x);
* TV-regularised least squares by torch Adam as the consumer.
Workload: 128^3 random ellipsoids, 120 views over 2*pi, 192^2 detector.
"""
import os

import numpy as np
import pytest
import torch

from cases import rel_l2

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def c5_operator(arith="exact", n=128, ns=120, nd=192):
    from ct2 import configs3D
    from ct2.operator import ConeBeamOperator
    R, SDD, dx, du = 500.0, 1000.0, 1.0, 1.6
    fv = configs3D.frame_vectors("circular_conebeam", {"radius": R, "source_to_detector": SDD},
                                 ns, 2 * np.pi, 0.0, 0)
    hw = n * dx / 2
    return ConeBeamOperator(fv, ns, nd, nd, du, du, -nd * du / 2, -nd * du / 2, (n, n, n),
                            (dx, dx, dx), (-hw, -hw, -hw), arith=arith)


def c5_volume(n=128, seed=20260105):
    from ct2 import phantoms
    v = phantoms.random_ellipsoids(n, n / 2, 40, seed)
    return torch.from_numpy(np.ascontiguousarray(v, np.float32)).cuda()


@pytest.fixture(scope="module")
def op():
    return c5_operator("exact")


def test_autograd_dot_product(op):
    from ct2.autograd import backproject, project
    g = torch.Generator(device="cpu").manual_seed(20260106)
    x = torch.randn(op.shape, generator=g).cuda().requires_grad_(True)
    y = torch.randn(op.sino_shape, generator=g).cuda().requires_grad_(True)
    Ax = project(x, op)
    lhs = (Ax.double() * y.double()).sum()
    Aty = backproject(y, op)
    rhs = (x.double() * Aty.double()).sum()
    assert abs(lhs.item() - rhs.item()) <= 1e-5 * abs(lhs.item())
    # the gradient of <Ax, y> w.r.t. x is A^T y, and w.r.t. y is Ax
    lhs.backward()
    assert rel_l2(x.grad.cpu().numpy(), Aty.detach().cpu().numpy()) < 1e-6
    assert rel_l2(y.grad.cpu().numpy(), Ax.detach().cpu().numpy()) < 1e-6


def test_gradient_matches_reference_backprojection():
    """x = vol/2, b = A vol: 1/2||Ax-b||^2 has gradient A^T(Ax-b) = -A^T(A vol)/2,
    and A^T(A vol) is the reference Fortran's backprojection of its own
    projection (fixture).  Halving is exact in FP32, so the residual equals
    -b/2 bit for bit."""
    from ct2.autograd import project
    from ct2.operator import ConeBeamOperator
    z = np.load(os.path.join(GOLD, "c1_reference.npz"))
    ns, nu, nv, du, dv, u0, v0 = z["geo"]
    dx, dy, dz, x0, y0, z0, nx, ny, nz = z["grid"]
    A = ConeBeamOperator(z["frame_vectors"], int(ns), int(nu), int(nv), du, dv, u0, v0,
                         (int(nx), int(ny), int(nz)), (dx, dy, dz), (x0, y0, z0), arith="exact")
    vol = torch.from_numpy(z["vol"]).cuda()
    b = torch.from_numpy(z["sino"]).cuda()
    x = (0.5 * vol).requires_grad_(True)
    r = project(x, A) - b
    assert torch.equal(r.detach(), -0.5 * b)
    loss = 0.5 * (r * r).sum(dtype=torch.float64)
    loss.backward()
    ref = -0.5 * z["bp"].astype(np.float64)
    assert rel_l2(x.grad.cpu().numpy(), ref) < 1e-6


def test_finite_differences_fp64(op):
    """Central differences with step 1e-3 in FP64.  The data term uses the
    operator's linearity, A(x +- h e_i) = Ax +- A(h e_i), with both
    projections formed on the device and the loss summed in FP64.  The TV
    term evaluates totalvar.f:35-62 at x +- h e_i on the 3x3x3 block around
    voxel i: every term of the sum that involves x_i lies in that block, so
    the block's difference is the whole sum's, and the block's small value
    keeps the function's Real*4 result precise enough."""
    from ct2.autograd import project, total_variation, tv_value_dev
    lam, eps, h = 1e-3, 1e-8, 1e-3
    rng = np.random.default_rng(20260107)
    x0 = torch.from_numpy(rng.uniform(0.2, 1.0, op.shape).astype(np.float32)).cuda()
    b = op.forward(c5_volume())
    x = x0.clone().requires_grad_(True)
    r = project(x, op) - b
    loss = 0.5 * (r * r).sum(dtype=torch.float64) + lam * total_variation(x, 0.0, eps)
    loss.backward()
    grad = x.grad.double().cpu().numpy()
    r64 = (op.forward(x0).double() - b.double())
    n = op.shape[0]
    idx = rng.integers(2, n - 2, size=(64, 3))
    fd = np.zeros(64)
    fd_tv = np.zeros(64)
    tv_grad = np.zeros(64)
    from ct2.autograd import tv_grad_dev
    gtv = tv_grad_dev(x0, 0.0, eps, normalizeq=0)[0].double().cpu().numpy()
    for q, (i, j, k) in enumerate(idx):
        e = torch.zeros_like(x0)
        e[i, j, k] = h
        c = op.forward(e).double()
        lp = 0.5 * ((r64 + c) ** 2).sum()
        lm = 0.5 * ((r64 - c) ** 2).sum()
        blk = (slice(i - 1, i + 2), slice(j - 1, j + 2), slice(k - 1, k + 2))
        tp = tv_value_dev((x0 + e)[blk].contiguous(), 0.0, eps).item()
        tm = tv_value_dev((x0 - e)[blk].contiguous(), 0.0, eps).item()
        fd_tv[q] = (tp - tm) / (2 * h)
        fd[q] = (lp - lm).item() / (2 * h) + lam * fd_tv[q]
        tv_grad[q] = gtv[i, j, k]
    g_at = grad[idx[:, 0], idx[:, 1], idx[:, 2]]
    assert np.max(np.abs(g_at - fd)) <= 1e-4 * np.max(np.abs(fd)) + 1e-6
    # the TV part on its own (it is 1e-3 of the total above)
    assert np.max(np.abs(tv_grad - fd_tv)) <= 2e-3 * np.max(np.abs(fd_tv))


def test_tv_least_squares_adam():
    """1/2||Ax-b||^2 + 1e-3*TV(x), torch Adam for 200 steps (fast arithmetic)."""
    from ct2.solvers import tv_least_squares
    A = c5_operator("fast")
    truth = c5_volume()
    b = A.forward(truth)
    x, losses = tv_least_squares(A, b, lam=1e-3, steps=200, lr=1e-2)
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.02 * losses[0]
    err0 = rel_l2(np.zeros(A.shape), truth.cpu().numpy())
    err = rel_l2(x.cpu().numpy(), truth.cpu().numpy())
    assert err < 0.5 * err0, (err, err0)
