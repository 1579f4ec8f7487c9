"""Pins of the SART composition in oracle/sart.py (no counterpart in the
reference; built from the pinned sinoproj / sinobackproject of
emil_code/view_proj3D.f:703-1595).

* Dense brute force: on a tiny grid the matrix A is assembled column by column
  from the forward projector, and the update is evaluated with explicit
  matrices (row sums, column sums over the subset's used rays, guards).
* Consistent data b = A x* are a fixed point of every subset's update.
* On a tiny system of full column rank, SART sweeps converge to x* (Censor &
  Elfving 2002: SART with 0 < lam < 2 converges to a solution of a
  consistent system).
* Guards: a ray that misses the volume and a voxel no used ray reaches leave
  the update finite and the voxel unchanged.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from oracle import sart as osart
from oracle import LINEAR, NEAREST
from ct2 import configs3D


def _case(n=(5, 4, 3), ns=6, nu=9, nv=7, du=1.5, dv=1.5, u_shift=0.0, seed=0, holes=True):
    fv = configs3D.frame_vectors("circular_conebeam", {"radius": 60.0, "source_to_detector": 120.0},
                                 ns, 2 * np.pi, 0.2, 0)
    nx, ny, nz = n
    geo = (ns, nu, nv, du, dv, -nu * du / 2 + u_shift, -nv * dv / 2)
    g2 = (1.0, 1.0, 1.0, -nx / 2, -ny / 2, -nz / 2, nx, ny, nz)
    rng = np.random.default_rng(seed)
    ind = (rng.random((ns, nu, nv)) > (0.1 if holes else -1)).astype(np.int8)
    return fv, geo, g2, ind, rng


def _dense(o, interp, ind, fv, geo, g2):
    nx, ny, nz = g2[-3:]
    cols = []
    for q in range(nx * ny * nz):
        e = np.zeros(nx * ny * nz)
        e[q] = 1
        cols.append(o.sinoproj(interp, 0, ind, fv, *geo, e.reshape(nx, ny, nz), *g2).ravel())
    return np.stack(cols, axis=1)


@pytest.fixture(scope="module")
def o64():
    return oracle.restated_raw("f64")


@pytest.mark.parametrize("interp", [LINEAR, NEAREST])
@pytest.mark.parametrize("view0,vstride", [(0, 1), (1, 3), (2, 2)])
def test_sart_matches_dense_matrices(o64, interp, view0, vstride):
    # narrow in u (corner voxels unseen), tall in v (outer rows miss the volume)
    fv, geo, g2, ind, rng = _case(n=(7, 6, 3), nu=2, nv=12, du=0.6)
    ns, nu, nv = geo[:3]
    A = _dense(o64, interp, ind, fv, geo, g2)
    x = rng.random(g2[-3:])
    b = rng.random((ns, nu, nv))
    views = osart.subset_views(ns, view0, 0, vstride)
    rows = np.zeros((ns, nu, nv), bool)
    rows[views] = True
    m = (rows & (ind == 1)).ravel()
    As = A[m]
    rs = As.sum(1)
    cs = As.T @ np.ones(m.sum())
    r = np.where(rs > 0, (b.ravel()[m] - As @ x.ravel()) / np.where(rs > 0, rs, 1), 0)
    want = x.ravel() + np.where(cs > 0, 0.7 * (As.T @ r) / np.where(cs > 0, cs, 1), 0)
    got = osart.sart_step(o64, interp, 0, 0.7, b, ind, fv, *geo, x, *g2, view0=view0,
                          vstride=vstride)
    assert (rs == 0).any() and (cs == 0).any(), "the case must exercise both guards"
    assert np.allclose(got.ravel(), want, rtol=1e-12, atol=1e-12)


def test_consistent_data_fixed_point(o64):
    fv, geo, g2, ind, rng = _case()
    x = rng.random(g2[-3:])
    b = o64.sinoproj(LINEAR, 0, ind, fv, *geo, x, *g2)
    for s in range(3):
        y = osart.sart_step(o64, LINEAR, 0, 1.0, b, ind, fv, *geo, x, *g2, view0=s, vstride=3)
        assert np.allclose(y, x, rtol=0, atol=1e-13)


def test_sart_converges_to_solution(o64):
    fv, geo, g2, ind, rng = _case(n=(4, 4, 4), ns=24, nu=12, nv=12, du=1.0, dv=1.0, holes=False)
    A = _dense(o64, LINEAR, ind, fv, geo, g2)
    assert np.linalg.matrix_rank(A) == A.shape[1], "the system must have full column rank"
    xs = rng.random(g2[-3:])
    b = o64.sinoproj(LINEAR, 0, ind, fv, *geo, xs, *g2)
    x = osart.sart_sweeps(o64, LINEAR, 0, 1.0, 200, 3, b, ind, fv, *geo, np.zeros(g2[-3:]), *g2)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-3


def test_guards_keep_unseen_voxels(o64):
    fv, geo, g2, ind, rng = _case(nu=4, nv=12, du=1.0, u_shift=2.0)
    ns, nu, nv = geo[:3]
    x = rng.random(g2[-3:])
    b = rng.random((ns, nu, nv))
    y = osart.sart_step(o64, LINEAR, 0, 1.0, b, ind, fv, *geo, x, *g2, view0=0, nviews=1)
    col = o64.sinobackproject(LINEAR, 0, (ind[:1] == 1).astype(float), ind[:1], fv[:1],
                              1, *geo[1:], *g2)
    assert np.isfinite(y).all()
    assert (col == 0).any()
    assert np.array_equal(y[col == 0], x[col == 0])
