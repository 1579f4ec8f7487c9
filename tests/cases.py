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


def quadric_objects(scale=1.0):
    """Analytic objects of every surface kind, sized for a volume of half
    width ~16*scale around the origin: a capped and tilted cylinder, a capped
    cone, an ellipsoid cut by a plane, a box of six planes, and two closed-form
    ellipsoids."""
    from ct2 import quadric as q
    s = scale
    cyl = q.gen_object(1.0 * s, -0.5 * s, 0.3 * s, att=0.8)
    cyl.add_surface(q.cylindrical(0., 0., 0., 0., 0., 0., ax=6. * s, ay=4. * s))
    cyl.add_surface(q.planar(0., 0., -7. * s))                        # z >= -7
    cyl.add_surface(q.planar(0., 0., 7. * s, 0., np.pi, 0.))          # z <= 7
    cyl.rotate(0.4, 0.9, -0.3)
    cone = q.gen_object(-2. * s, 1. * s, -6. * s, att=1.3)
    cone.add_surface(q.conical(0., 0., 0., 0., 0., 0., ax=0.7, ay=0.5))
    cone.add_surface(q.planar(0., 0., 11. * s, 0., np.pi, 0.))        # z <= 11
    cone.rotate(0.2, -0.35, 0.0)
    cut = q.gen_object(2. * s, 2. * s, -1. * s, att=-0.4)
    cut.add_surface(q.planar(0.5 * s, 0., 0., 0.3, 1.2, 0.))
    cut.add_surface(q.ellipsoidal(0., 0., 0., 0.5, 0.2, 0.1, ax=7. * s, ay=5. * s, az=9. * s))
    box = q.gen_object(0., 0., 0., att=0.25)
    for axis, sgn in ((0, 1), (0, -1), (1, 1), (1, -1), (2, 1), (2, -1)):
        # frame z along +-axis: plane through -9s along the normal
        beta, alpha = {(2, 1): (0., 0.), (2, -1): (np.pi, 0.), (0, 1): (np.pi / 2, 0.),
                       (0, -1): (-np.pi / 2, 0.), (1, 1): (np.pi / 2, np.pi / 2),
                       (1, -1): (-np.pi / 2, np.pi / 2)}[(axis, sgn)]
        c = [0., 0., 0.]
        c[axis] = -sgn * 9. * s
        box.add_surface(q.planar(*c, alpha, beta, 0.))
    e1 = q.ellipsoid(1. * s, -2. * s, 3. * s, att=0.9, ax=8. * s, ay=5. * s, az=11. * s,
                     alpha=0.3, beta=1.1)
    e2 = q.ellipsoid(-4. * s, 3. * s, -2. * s, att=0.35, ax=3. * s, ay=3. * s, az=3. * s)
    return [cyl, cone, cut, box, e1, e2]


def f2py_surface_args(obj):
    """gen_shape_proj's arguments after att for a gen_object, packed as the
    reference's gen_object.project_to packs them (tomo3D.py:2400-2530)."""
    cols = [[] for _ in range(12)]
    ax = {"ellipsoid": ([1.], [1.], [1.]), "cylinder": ([1.], [1.]), "cone": ([1.], [1.])}
    counts = {"plane": 1, "ellipsoid": 1, "cylinder": 1, "cone": 1}
    for label in ("plane", "ellipsoid", "cylinder", "cone"):
        for s in obj.surfaces:
            if s.label != label:
                continue
            for c, v in enumerate((s.xu1, s.xu2, s.xu3, s.yu1, s.yu2, s.yu3, s.zu1, s.zu2, s.zu3,
                                   s.xc + obj.x0, s.yc + obj.y0, s.zc + obj.z0)):
                cols[c].append(v)
            if label != "plane":
                for lst, v in zip(ax[label], s.axes()):
                    lst.append(v)
            counts[label] += 1
    return (obj.nsurf, *cols, counts["plane"], counts["ellipsoid"], *ax["ellipsoid"],
            counts["cylinder"], *ax["cylinder"], counts["cone"], *ax["cone"])


def ellipsoid_args(e):
    return (e.ax, e.ay, e.az, e.x0, e.y0, e.z0, e.alpha, e.beta, e.att)
