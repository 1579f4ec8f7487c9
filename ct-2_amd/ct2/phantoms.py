"""Seeded voxel phantoms for the benchmark workloads.

The reference's own phantom classes (emil_code/tomo3D.py:1507-2894,
emil_code/phantoms/*.py) are data generators outside the reconstruction
path; parity tests feed the same voxel array to every implementation, so the
phantom only has to be deterministic.  Ellipsoids are axis-aligned unless a
rotation about z is given; a voxel is inside when its centre satisfies
sum(((x-c)/a)^2) <= 1 (emil_code/tomo3D.py:2793).
"""
from __future__ import annotations

import numpy as np


def voxel_centres(n, lo, d):
    return lo + (np.arange(n) + 0.5) * d


def add_ellipsoids(vol, grid, ells):
    """vol[nx,ny,nz] += rho inside each ellipsoid (cx,cy,cz,ax,ay,az,rho,phi_deg).
    grid = (x0,y0,z0,dx,dy,dz) in the same length unit as the ellipsoids."""
    nx, ny, nz = vol.shape
    x0, y0, z0, dx, dy, dz = grid
    xs = voxel_centres(nx, x0, dx)
    ys = voxel_centres(ny, y0, dy)
    zs = voxel_centres(nz, z0, dz)
    for (cx, cy, cz, ax, ay, az, rho, phi) in ells:
        # restrict to the bounding box for speed
        r = max(ax, ay, az)
        ix = np.nonzero(np.abs(xs - cx) <= r + dx)[0]
        iy = np.nonzero(np.abs(ys - cy) <= r + dy)[0]
        iz = np.nonzero(np.abs(zs - cz) <= az + dz)[0]
        if ix.size == 0 or iy.size == 0 or iz.size == 0:
            continue
        X = xs[ix][:, None, None] - cx
        Y = ys[iy][None, :, None] - cy
        Z = zs[iz][None, None, :] - cz
        if phi:
            t = np.deg2rad(phi)
            Xr = X * np.cos(t) + Y * np.sin(t)
            Yr = -X * np.sin(t) + Y * np.cos(t)
        else:
            Xr, Yr = X, Y
        inside = (Xr / ax) ** 2 + (Yr / ay) ** 2 + (Z / az) ** 2 <= 1.0
        sub = vol[ix[0]:ix[-1] + 1, iy[0]:iy[-1] + 1, iz[0]:iz[-1] + 1]
        sub += (rho * inside).astype(vol.dtype)
    return vol


# 10-ellipsoid 3D Shepp-Logan in normalised [-1,1]^3 (Kak & Slaney geometry,
# Toft's modified densities): cx, cy, cz, ax, ay, az, rho, phi (deg about z)
SHEPP_LOGAN_10 = [
    (0.0, 0.0, 0.0, 0.69, 0.92, 0.81, 1.0, 0.0),
    (0.0, -0.0184, 0.0, 0.6624, 0.874, 0.78, -0.8, 0.0),
    (0.22, 0.0, 0.0, 0.11, 0.31, 0.22, -0.2, -18.0),
    (-0.22, 0.0, 0.0, 0.16, 0.41, 0.28, -0.2, 18.0),
    (0.0, 0.35, -0.15, 0.21, 0.25, 0.41, 0.1, 0.0),
    (0.0, 0.1, 0.25, 0.046, 0.046, 0.05, 0.1, 0.0),
    (0.0, -0.1, 0.25, 0.046, 0.046, 0.05, 0.1, 0.0),
    (-0.08, -0.605, 0.0, 0.046, 0.023, 0.05, 0.1, 0.0),
    (0.0, -0.606, 0.0, 0.023, 0.023, 0.02, 0.1, 0.0),
    (0.06, -0.605, 0.0, 0.023, 0.046, 0.02, 0.1, 0.0),
]


def shepp_logan(n, half_width):
    """n^3 float32 Shepp-Logan occupying [-half_width, half_width]^3."""
    vol = np.zeros((n, n, n), np.float64)
    d = 2.0 * half_width / n
    ells = [(cx * half_width, cy * half_width, cz * half_width,
             ax * half_width, ay * half_width, az * half_width, rho, phi)
            for (cx, cy, cz, ax, ay, az, rho, phi) in SHEPP_LOGAN_10]
    add_ellipsoids(vol, (-half_width,) * 3 + (d,) * 3, ells)
    return vol.astype(np.float32)


def random_ellipsoids(n, half_width, count, seed, centre=0.7, axes=(0.03, 0.4),
                      dens=(0.05, 1.0)):
    """n^3 float32 volume of `count` additive axis-aligned ellipsoids with
    centres U[-centre,centre]^3, semi-axes U[axes], densities U[dens], in
    normalised [-1,1]^3 scaled by half_width."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-centre, centre, (count, 3))
    a = rng.uniform(axes[0], axes[1], (count, 3))
    rho = rng.uniform(dens[0], dens[1], count)
    ells = [(*(c[i] * half_width), *(a[i] * half_width), rho[i], 0.0) for i in range(count)]
    vol = np.zeros((n, n, n), np.float64)
    d = 2.0 * half_width / n
    add_ellipsoids(vol, (-half_width,) * 3 + (d,) * 3, ells)
    return vol.astype(np.float32)


def breast_like(shape, voxel, seed, n_background=200, n_calc=50):
    """Tomosynthesis phantom: fatty-glandular background of ellipsoids
    (density 0.2-0.35) plus microcalcification spheres of radius 0.1-0.3
    (length unit of `voxel`) at density 2.0.  shape=(nx,ny,nz), voxel=(dx,dy,dz);
    the volume is centred on x and y and spans z in [0, nz*dz)."""
    nx, ny, nz = shape
    dx, dy, dz = voxel
    X, Y, Z = nx * dx, ny * dy, nz * dz
    rng = np.random.default_rng(seed)
    vol = np.zeros(shape, np.float32)
    grid = (-X / 2, -Y / 2, 0.0, dx, dy, dz)
    ells = []
    for _ in range(n_background):
        c = rng.uniform([-0.45 * X, -0.45 * Y, 0.1 * Z], [0.45 * X, 0.45 * Y, 0.9 * Z])
        a = rng.uniform([0.03 * X, 0.03 * Y, 0.1 * Z], [0.15 * X, 0.15 * Y, 0.4 * Z])
        ells.append((*c, *a, rng.uniform(0.2, 0.35), 0.0))
    for _ in range(n_calc):
        c = rng.uniform([-0.4 * X, -0.4 * Y, 0.2 * Z], [0.4 * X, 0.4 * Y, 0.8 * Z])
        r = rng.uniform(0.1, 0.3)
        ells.append((*c, r, r, r, 2.0, 0.0))
    add_ellipsoids(vol, grid, ells)
    return vol


def random_ellipsoids_device(shape, voxel, count, rng, device="cuda"):
    """Additive random axis-aligned ellipsoids rasterised into a float32
    device tensor: centres U[-0.7,0.7], semi-axes U[0.03,0.4], densities
    U[0.05,1.0], lengths in units of the half-width of the volume's largest
    side; ``rng`` (a numpy Generator) draws the table, so a seed fixes it."""
    import torch
    hw = max(shape) * voxel / 2
    ax = [(torch.arange(m, device=device, dtype=torch.float32) + 0.5 - m / 2) * voxel
          for m in shape]
    X, Y, Z = ax[0][:, None, None], ax[1][None, :, None], ax[2][None, None, :]
    vol = torch.zeros(shape, dtype=torch.float32, device=device)
    c = rng.uniform(-0.7, 0.7, (count, 3)) * hw
    a = rng.uniform(0.03, 0.4, (count, 3)) * hw
    rho = rng.uniform(0.05, 1.0, count)
    for i in range(count):
        q = ((X - c[i, 0]) / a[i, 0]) ** 2 + ((Y - c[i, 1]) / a[i, 1]) ** 2 + \
            ((Z - c[i, 2]) / a[i, 2]) ** 2
        vol += float(rho[i]) * (q <= 1.0).float()
    return vol
