"""Closed-form pins of the FDK restatement (oracle/fdk.py, restating
emil_code/FDK.py:10-208 and setFreqrange.py:3-10, called as
recon3D.py:53-131 calls them).

Every check derives its expected value from geometry or mathematics that the
restatement does not use:
  * the detector position (u, v) and magnification of a voxel come from a
    vector ray-plane intersection in the frame of configs3D.py:11-49
    (source R(cos s, sin s, 0), detector centre (R-SDD)(cos s, sin s, 0),
    e_u = (-sin s, cos s, 0), e_v = z), not from the zoom formula;
  * an all-ones sinogram backprojects to the angular integral
    pi SDD^2 R / (R^2 - r^2)^(3/2) (the trapezoid sum of a periodic analytic
    function over ns equispaced views is exact to rounding here);
  * bilinear interpolation reproduces affine detector functions exactly, so
    a + b u + c v backprojects to pi/ns sum_s (a + b u + c v) M^2;
  * FDK of a ball's exact projections reconstructs its density scaled by
    SDD/(R+SDD), the factor the reference's weight call with D = SDD
    (recon3D.py:65-67) gives against the correctly weighted FDK;
  * the filter is a circular convolution with the Kak-Slaney kernel
    (Kak & Slaney, Principles of Computerized Tomographic Imaging, eq. 3.61),
    evaluated as a direct sum, including the wrap-around when
    npad < 2n - 1 (FDK.py:36 rounds log2(2n-1));
  * the Hann window divided out of the filter equals 0.5 + 0.5 cos(pi f/f_c)
    on numpy.fft.fftfreq's grid.
test_mutants_are_caught plants mistakes in a copy of the restatement's
source and requires every mutant to fail at least one check.
"""
from __future__ import annotations

import inspect
import types

import numpy as np
import pytest

from oracle import fdk as ofdk

R, SDD = 100.0, 200.0


def _geometry(P, s):
    """Detector coordinates (u, v) and magnification M of point P for view
    angle s, by intersecting the ray from the source through P with the
    detector plane."""
    S = np.array([R * np.cos(s), R * np.sin(s), 0.0])
    D = (R - SDD) * np.array([np.cos(s), np.sin(s), 0.0])
    eu = np.array([-np.sin(s), np.cos(s), 0.0])
    ev = np.array([0.0, 0.0, 1.0])
    ew = np.cross(eu, ev)
    t = np.dot(D - S, ew) / np.dot(P - S, ew)
    X = S + t * (P - S)
    return np.dot(X - D, eu), np.dot(X - D, ev), t


def check_weight(m):
    """The pre-weight with the reference's call (D = SDD): R/(R+SDD) on the
    axis, times the cosine of the ray angle to the detector normal."""
    u = np.array([[0.0, 13.0], [-7.0, 30.0]])
    v = np.array([[0.0, -4.0], [22.0, 9.0]])
    w = m.fdk_sino_weight3D(R, SDD, 1, u, v, np.ones((1, 2, 2)))[0]
    L = R + SDD
    cosg = np.array([[np.dot([a, b, L], [0, 0, 1]) / np.linalg.norm([a, b, L])
                      for a, b in zip(ur, vr)] for ur, vr in zip(u, v)])
    return np.abs(w - R / L * cosg).max() < 1e-14


def check_ones(m):
    """All-ones sinogram over a full scan: pi SDD^2 R / (R^2 - r^2)^(3/2)."""
    ns, nu, nv, du = 72, 160, 160, 0.5
    u0 = -nu * du / 2
    vs = np.arange(ns) * 2 * np.pi / ns + 0.1
    vx = np.array([0.0, 6.0, -9.5, 3.0])
    vy = np.array([0.0, -2.0, 4.0, 10.0])
    vz = np.array([0.0, 7.0])
    rec = m.fdk_backproj_linear3D(ns, nu, nv, du, du, u0, u0, R, SDD, 0.5, 0.5, vs, vx, vy, vz,
                                  np.ones((ns, nu, nv)))
    r2 = vx ** 2 + vy ** 2
    want = np.pi * SDD ** 2 * R / (R ** 2 - r2) ** 1.5
    return np.allclose(rec, want[:, None], rtol=1e-10)


def check_affine(m):
    """a + b u + c v on the detector, four views, three slices, points whose
    projections sit inside and outside the detector."""
    ns, nu, nv = 4, 90, 70
    du, dv = 0.5, 0.4
    u0, v0 = -nu * du / 2 + 1.3, -nv * dv / 2 - 0.7
    a, b, c = 1.0, 0.75, -1.6
    vs = np.array([0.3, 1.4, 2.9, 4.4])
    uc = u0 + (np.arange(nu) + 0.5) * du
    vc = v0 + (np.arange(nv) + 0.5) * dv
    sino = np.broadcast_to(a + b * uc[:, None] + c * vc[None, :], (ns, nu, nv)).copy()
    vx = np.array([0.0, 5.0, -8.0, 2.5, 19.0])
    vy = np.array([0.0, -3.0, 6.0, 9.0, -1.0])
    vz = np.array([-6.0, 0.0, 4.25])
    rec = m.fdk_backproj_linear3D(ns, nu, nv, du, dv, u0, v0, R, SDD, 0.5, 0.5, vs, vx, vy, vz,
                                  sino)
    want = np.zeros((vx.size, vz.size))
    for q in range(vx.size):
        for z in range(vz.size):
            for s in vs:
                u, v, M = _geometry(np.array([vx[q], vy[q], vz[z]]), s)
                lu = np.floor((u - u0) / du - 0.5)
                lv = np.floor((v - v0) / dv - 0.5)
                if 0 <= lu and lu + 1 < nu and 0 <= lv and lv + 1 < nv:
                    want[q, z] += (a + b * u + c * v) * M * M
    want *= np.pi / ns
    return np.allclose(rec, want, rtol=1e-9, atol=1e-9) and np.abs(want).min() >= 0.0


def check_ball(m):
    """FDK (weight with D = SDD, ramp along u, backprojection) of an
    off-centre ball's exact chords: density x SDD/(R+SDD) inside, ~0 away."""
    ns, nu, nv, du = 120, 112, 112, 0.5
    rad, rho = 6.0, 1.0
    cen = np.array([4.0, -3.0, 2.0])
    u0 = -nu * du / 2
    uc = u0 + (np.arange(nu) + 0.5) * du
    vs = np.arange(ns) * 2 * np.pi / ns
    sino = np.zeros((ns, nu, nu))
    for p, s in enumerate(vs):
        S = np.array([R * np.cos(s), R * np.sin(s), 0.0])
        D = (R - SDD) * np.array([np.cos(s), np.sin(s), 0.0])
        eu = np.array([-np.sin(s), np.cos(s), 0.0])
        B = D + uc[:, None, None] * eu + uc[None, :, None] * np.array([0.0, 0.0, 1.0])
        d = B - S
        d /= np.linalg.norm(d, axis=2, keepdims=True)
        sc = cen - S
        tca = (sc * d).sum(2)
        sino[p] = rho * 2 * np.sqrt(np.clip(rad ** 2 - (sc ** 2).sum() + tca ** 2, 0, None))
    vv, vu = np.meshgrid(uc, uc)
    f = m.fdk_sino_filter3D(du, 1, m.fdk_sino_weight3D(R, SDD, ns, vu, vv, sino))
    n, dx = 28, 0.5
    x1 = -n * dx / 2 + (np.arange(n) + 0.5) * dx
    vxg, vyg = np.meshgrid(x1, x1)
    vz = np.array([2.0, 0.5])
    rec = m.fdk_backproj_linear3D(ns, nu, nu, du, du, u0, u0, R, SDD, 0.5, 0.5, vs,
                                  vxg.flatten(), vyg.flatten(), vz, f)
    cube = ofdk.recon_cube(rec, n, n, vz.size)
    X, Y = np.meshgrid(x1, x1, indexing="ij")
    ok = True
    for q, z in enumerate(vz):
        r2 = (X - cen[0]) ** 2 + (Y - cen[1]) ** 2 + (z - cen[2]) ** 2
        inner = cube[:, :, q][r2 < (0.6 * rad) ** 2]
        outer = cube[:, :, q][r2 > (1.5 * rad) ** 2]
        ok &= abs(inner.mean() / (rho * SDD / (R + SDD)) - 1) < 0.01
        ok &= inner.std() < 0.01 and np.abs(outer).max() < 0.06
    return bool(ok)


def _kak_slaney(k, dp):
    k = np.asarray(k)
    return np.where(k == 0, 1 / (4 * dp ** 2),
                    np.where(k % 2 != 0, -1 / (np.pi * np.where(k == 0, 1, k) * dp) ** 2, 0.0))


def check_filter_circular(m):
    """Filtering along u and along v equals du x the circular convolution of
    the zero-padded line with the Kak-Slaney kernel h(k), k in [-N/2, N/2),
    period N = npad; n = 40 pads to 64 < 79, so the wrap-around is live."""
    ok = True
    for n, dim, dp in ((40, 1, 0.7), (33, 2, 1.3)):
        N = int(2 ** np.round(np.log2(2 * n - 1)))
        x = np.random.default_rng(n).normal(size=(2, n, 3) if dim == 1 else (2, 3, n))
        y = m.fdk_sino_filter3D(dp, dim, x)
        i = np.arange(n)
        k = (i[:, None] - i[None, :] + N // 2) % N - N // 2
        H = dp * _kak_slaney(k, dp)
        want = np.einsum("im,amb->aib", H, x) if dim == 1 else np.einsum("im,abm->abi", H, x)
        ok &= np.allclose(y, want, rtol=1e-9, atol=1e-12)
    return bool(ok)


def check_hann(m):
    """Hann-windowed filter / plain ramp = 0.5 + 0.5 cos(pi f / f_c) on the
    fftfreq grid, f_c = L/(2 dp); for L < 1 zero beyond the grid frequencies
    nearest to +-f_c."""
    ok = True
    for n, dp, L in ((64, 0.8, 1.0), (50, 1.1, 0.5), (64, 0.8, 0.75)):
        npad, ramp = m.ramp_filter(dp, n)
        _, hann = m.ramp_filter(dp, n, "Hann", L)
        f = np.fft.fftfreq(npad, dp)
        fc = L / (2 * dp)
        win = 0.5 + 0.5 * np.cos(np.pi * f / fc)
        if L < 1:
            near = np.abs(np.abs(f) - fc)
            win = np.where(np.abs(f) <= f[np.argmin(near)] * np.sign(f[np.argmin(near)]) + 1e-12,
                           win, 0.0)
        ok &= np.allclose(hann, ramp * win, rtol=1e-12, atol=1e-12)
    return bool(ok)


CHECKS = [check_weight, check_ones, check_affine, check_ball, check_filter_circular, check_hann]


@pytest.mark.parametrize("check", CHECKS, ids=lambda c: c.__name__)
def test_restatement_passes(check):
    assert check(ofdk)


# Each mutant is (description, original text, replacement) in oracle/fdk.py.
MUTANTS = [
    ("zoom instead of zoom^2", "val * zoomfac[m] * zoomfac[m]", "val * zoomfac[m]"),
    ("no pi/ns", "return np.pi / ns * recon", "return recon"),
    ("u sign flipped", "uinlen = (-vx * np.sin(vs[ii]) + vy * np.cos(vs[ii]))",
     "uinlen = (vx * np.sin(vs[ii]) - vy * np.cos(vs[ii]))"),
    ("v without magnification", "vinlen = vz[jj] * zoomfac", "vinlen = vz[jj] + 0 * zoomfac"),
    ("u and v swapped", "uinlen = (-vx * np.sin(vs[ii]) + vy * np.cos(vs[ii])) * zoomfac\n"
     "            vinlen = vz[jj] * zoomfac",
     "vinlen = (-vx * np.sin(vs[ii]) + vy * np.cos(vs[ii])) * zoomfac\n"
     "            uinlen = vz[jj] * zoomfac"),
    ("D = SDD in the backprojection", "D = source2detector - radius", "D = source2detector"),
    ("weight without F", "weight = F / np.sqrt((F + D) ** 2", "weight = F / np.sqrt((D) ** 2"),
    ("no ifftshift of h", "hn = np.fft.ifftshift(hn)", "hn = hn"),
    ("no dp factor", "ramp = dp * Hk", "ramp = Hk"),
    ("Hann cutoff doubled", "cutoff_freq = 1 / (2 * dp) * L", "cutoff_freq = 1 / dp * L"),
    ("bin offset dropped", "lu = np.floor((uinlen - u0) / du - uoffset)",
     "lu = np.floor((uinlen - u0) / du)"),
]


def _mutant(old, new):
    src = inspect.getsource(ofdk)
    assert src.count(old) == 1, f"mutation anchor not found once: {old!r}"
    mod = types.ModuleType("fdk_mutant")
    exec(compile(src.replace(old, new), "fdk_mutant", "exec"), mod.__dict__)
    return mod


@pytest.mark.parametrize("desc,old,new", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_mutants_are_caught(desc, old, new):
    mod = _mutant(old, new)
    failed = []
    for c in CHECKS:
        try:
            ok = c(mod)
        except Exception:
            ok = False
        if not ok:
            failed.append(c.__name__)
    assert failed, f"mutant '{desc}' passes every check"
