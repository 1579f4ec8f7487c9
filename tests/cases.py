"""Small seeded problem instances shared by the CPU and GPU parity tests."""
from __future__ import annotations

import numpy as np

from ct2 import configs3D, phantoms


def cone_case(n=32, ns=12, nu=40, nv=36, seed=0, phantom="shepp", R=500.0, sdd=1000.0,
              du=1.5, dv=1.5, dx=1.0, holes=True):
    """Circular cone beam, s_include_endpoints=0 over 2*pi; voxel dx; detector
    centred.  Returns (ind, fv, geo, vol, g2) in the f2py argument order."""
    fv = configs3D.frame_vectors("circular_conebeam", {"radius": R, "source_to_detector": sdd},
                                 ns, 2 * np.pi, 0.0, 0)
    hw = n * dx / 2
    if phantom == "shepp":
        vol = phantoms.shepp_logan(n, hw)
    else:
        vol = phantoms.random_ellipsoids(n, hw, 12, seed)
    ind = np.ones((ns, nu, nv), np.int8)
    if holes:
        rng = np.random.default_rng(seed + 100)
        ind[rng.random((ns, nu, nv)) < 0.02] = 0
    geo = (ns, nu, nv, du, dv, -nu * du / 2, -nv * dv / 2)
    g2 = (dx, dx, dx, -hw, -hw, -hw, n, n, n)
    return ind, fv, geo, vol, g2


def tomo_case(seed=0, ns=7, nu=48, nv=40, nx=40, ny=32, nz=8):
    fv = configs3D.frame_vectors("tomosynthesis", {"radius": 650.0, "center_to_detector": 20.0},
                                 ns, np.deg2rad(50), np.deg2rad(65), 1)
    rng = np.random.default_rng(seed)
    vol = rng.random((nx, ny, nz)).astype(np.float32)
    geo = (ns, nu, nv, 1.0, 1.0, -nu / 2, -nv / 2)
    g2 = (0.8, 0.8, 2.0, -nx * 0.4, -ny * 0.4, 0.0, nx, ny, nz)
    ind = np.ones((ns, nu, nv), np.int8)
    return ind, fv, geo, vol, g2


def ray_case(seed=0, nr=3000, shape=(24, 28, 20)):
    rng = np.random.default_rng(seed)
    rv = np.zeros((nr, 6), np.float32)
    for r in range(nr):
        th, ph = rng.uniform(0.2, 2.9), rng.uniform(0, 2 * np.pi)
        rv[r] = configs3D.lino_spherical_flatpanel(
            (th, ph, rng.uniform(-30, 30), rng.uniform(-30, 30)),
            {"source_radius": 300.0, "source_detector": 600.0})
    nx, ny, nz = shape
    vol = rng.random(shape).astype(np.float32)
    g3 = (1.0, 1.0, 1.0, -nx / 2, -ny / 2, -nz / 2, nx, ny, nz)
    ind = (rng.random(nr) > 0.1).astype(np.int8)
    return ind, rv, nr, vol, g3


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
