"""Analytic phantoms: quadric objects, their exact projections and their voxel
embeddings, computed by ``ct2_phantom_project``, ``ct2_phantom_project_rays``
and ``ct2_phantom_embed`` (ct2_phantom.hip).

The classes mirror the reference's phantom model (emil_code/tomo3D.py:1507-2894):
``phantom3D`` is a list of additive components; a component is an
``ellipsoid`` (closed-form chord, view_proj3D.f:483-603) or a ``gen_object``,
the intersection of ``planar``, ``ellipsoidal``, ``cylindrical`` and
``conical`` surfaces (view_proj3D.f:175-455).  Where the reference makes one
Fortran call per component and view, ``phantom3D.project_to`` makes one
launch for the whole phantom over every view; the per-bin sums are added in
the reference's component order, so the sinogram is the same.

The surface frames follow tomo3D.py:1707-1786: a surface made with angles
(alpha, beta, gamma) has the frame R = Rz(alpha) Ry(beta) Rz(gamma) applied to
the unit vectors, with Rz(t): (x, y, z) -> (x cos t - y sin t, x sin t + y cos t, z)
and Ry(t): (x, y, z) -> (x cos t + z sin t, y, -x sin t + z cos t).
"""
from __future__ import annotations

import ctypes as C
from math import cos, sin

import numpy as np
import torch

from . import _dev
from ._lib import call

CT2_SHAPE_ELLIPSOID, CT2_SHAPE_GENERAL = 0, 1
CT2_SURF_PLANE, CT2_SURF_ELLIPSOID, CT2_SURF_CYLINDER, CT2_SURF_CONE = 0, 1, 2, 3
CT2_SURF_ANGLES = 16   # voxel values only: frame from angles, as framer
_KIND_OF_LABEL = {"plane": CT2_SURF_PLANE, "ellipsoid": CT2_SURF_ELLIPSOID,
                  "cylinder": CT2_SURF_CYLINDER, "cone": CT2_SURF_CONE}


class ct2_surface(C.Structure):
    _fields_ = [("kind", C.c_int), ("ax", C.c_float), ("ay", C.c_float), ("az", C.c_float),
                ("xu", C.c_float * 3), ("yu", C.c_float * 3), ("zu", C.c_float * 3),
                ("c", C.c_float * 3)]


class ct2_shape(C.Structure):
    _fields_ = [("kind", C.c_int), ("att", C.c_float), ("first", C.c_int), ("nsurf", C.c_int),
                ("ax", C.c_float), ("ay", C.c_float), ("az", C.c_float),
                ("x0", C.c_float), ("y0", C.c_float), ("z0", C.c_float),
                ("alpha", C.c_float), ("beta", C.c_float)]


def _rotate(v, alpha, beta, gamma):
    """R v with R = Rz(alpha) Ry(beta) Rz(gamma) (tomo3D.py:1721-1733)."""
    x, y, z = v
    xp, yp, zp = x * cos(gamma) - y * sin(gamma), x * sin(gamma) + y * cos(gamma), z
    xpp, ypp, zpp = xp * cos(beta) + zp * sin(beta), yp, -xp * sin(beta) + zp * cos(beta)
    return (xpp * cos(alpha) - ypp * sin(alpha), ypp * cos(alpha) + xpp * sin(alpha), zpp)


# ---------------------------------------------------------------- tables ----
class Tables:
    """Host shape and surface tables in the C-ABI layout."""

    def __init__(self):
        self.shapes = []
        self.surfaces = []

    def add_ellipsoid(self, att, ax, ay, az, x0, y0, z0, alpha=0.0, beta=0.0):
        self.shapes.append(ct2_shape(CT2_SHAPE_ELLIPSOID, att, 0, 0, ax, ay, az, x0, y0, z0,
                                     alpha, beta))

    def add_general(self, att, surfaces):
        """surfaces: iterable of (kind, (ax, ay, az), xu, yu, zu, centre)."""
        first = len(self.surfaces)
        for kind, axes, xu, yu, zu, c in surfaces:
            self.surfaces.append(ct2_surface(kind, *axes, (C.c_float * 3)(*xu),
                                             (C.c_float * 3)(*yu), (C.c_float * 3)(*zu),
                                             (C.c_float * 3)(*c)))
        self.shapes.append(ct2_shape(CT2_SHAPE_GENERAL, att, first, len(self.surfaces) - first))

    def arrays(self):
        sh = (ct2_shape * max(len(self.shapes), 1))(*self.shapes)
        sf = (ct2_surface * max(len(self.surfaces), 1))(*self.surfaces)
        return sh, len(self.shapes), sf, len(self.surfaces)


def tables_from_f2py(att, nsurf, xu1, xu2, xu3, yu1, yu2, yu3, zu1, zu2, zu3, x0, y0, z0,
                     nplane, nellipsoid, ax_ell, ay_ell, az_ell, ncyl, ax_cyl, ay_cyl,
                     ncone, ax_cone, ay_cone):
    """One general object from gen_shape_proj's arguments (view_proj3D.f:175-189):
    the surface arrays hold planes, then ellipsoids, cylinders and cones, and
    each count and semi-axis array carries a leading dummy entry."""
    cols = [np.asarray(a, np.float64).reshape(-1) for a in
            (xu1, xu2, xu3, yu1, yu2, yu3, zu1, zu2, zu3, x0, y0, z0)]
    nsurf = int(nsurf)
    counts = [int(nplane) - 1, int(nellipsoid) - 1, int(ncyl) - 1, int(ncone) - 1]
    if min(counts) < 0 or sum(counts) > nsurf or any(c.size < nsurf for c in cols):
        raise ValueError("inconsistent surface counts for a general object")
    axes = {CT2_SURF_ELLIPSOID: (ax_ell, ay_ell, az_ell), CT2_SURF_CYLINDER: (ax_cyl, ay_cyl),
            CT2_SURF_CONE: (ax_cone, ay_cone)}
    rows, s = [], 0
    for kind, cnt in enumerate(counts):
        for q in range(1, cnt + 1):
            a = [float(np.asarray(v).reshape(-1)[q]) for v in axes.get(kind, ())]
            a += [0.0] * (3 - len(a))
            u = [c[s] for c in cols]
            rows.append((kind, a, u[0:3], u[3:6], u[6:9], u[9:12]))
            s += 1
    t = Tables()
    t.add_general(att, rows)
    return t


def tables_from_f2py_angles(att, nsurf, alpha, beta, gam, x0, y0, z0, nplane, nellipsoid,
                            ax_ell, ay_ell, az_ell, ncyl, ax_cyl, ay_cyl, ncone, ax_cone, ay_cone):
    """One general object from gen_obj_voxval's arguments
    (tomo3DfortranAux.f:311-346): surface frames given by the angles alpha,
    beta, gam and the centres, taken by framer (tomo3DfortranAux.f:281-308)."""
    cols = [np.asarray(a, np.float64).reshape(-1) for a in (alpha, beta, gam, x0, y0, z0)]
    nsurf = int(nsurf)
    counts = [int(nplane) - 1, int(nellipsoid) - 1, int(ncyl) - 1, int(ncone) - 1]
    if min(counts) < 0 or sum(counts) > nsurf or any(c.size < nsurf for c in cols):
        raise ValueError("inconsistent surface counts for a general object")
    axes = {CT2_SURF_ELLIPSOID: (ax_ell, ay_ell, az_ell), CT2_SURF_CYLINDER: (ax_cyl, ay_cyl),
            CT2_SURF_CONE: (ax_cone, ay_cone)}
    rows, s = [], 0
    for kind, cnt in enumerate(counts):
        for q in range(1, cnt + 1):
            a = [float(np.asarray(v).reshape(-1)[q]) for v in axes.get(kind, ())]
            a += [0.0] * (3 - len(a))
            u = [c[s] for c in cols]
            rows.append((kind | CT2_SURF_ANGLES, a, u[0:3], (0., 0., 0.), (0., 0., 0.), u[3:6]))
            s += 1
    t = Tables()
    t.add_general(att, rows)
    return t


# ------------------------------------------------------------ device calls ----
def project_views(tables, sino, ind, fv, ns, nu, nv, du, dv, u0, v0, accumulate):
    """sino: contiguous float32 device tensor [ns, nu, nv] (written in place)."""
    sh, nsh, sf, nsf = tables.arrays()
    call("ct2_phantom_project", C.cast(sh, C.c_void_p), nsh, C.cast(sf, C.c_void_p), nsf,
         _dev.ptr(sino), _dev.ptr(ind), _dev.ptr(fv), ns, nu, nv, du, dv, u0, v0,
         int(accumulate), _dev.stream_handle())


def project_rays(tables, sino, ind, rv, nr, accumulate):
    sh, nsh, sf, nsf = tables.arrays()
    call("ct2_phantom_project_rays", C.cast(sh, C.c_void_p), nsh, C.cast(sf, C.c_void_p), nsf,
         _dev.ptr(sino), _dev.ptr(ind), _dev.ptr(rv), nr, int(accumulate), _dev.stream_handle())


def embed(tables, vol, nx, ny, nz, dx, dy, dz, x0, y0, z0, accumulate):
    sh, nsh, sf, nsf = tables.arrays()
    call("ct2_phantom_embed", C.cast(sh, C.c_void_p), nsh, C.cast(sf, C.c_void_p), nsf,
         _dev.ptr(vol), nx, ny, nz, dx, dy, dz, x0, y0, z0, int(accumulate),
         _dev.stream_handle())


def voxval_at(tables, x, y, z):
    """The phantom's voxel value at points (x, y, z), broadcast together, in
    one ct2_phantom_points launch: a float for scalar inputs, else a float32
    array of the broadcast shape."""
    b = np.broadcast_arrays(*(np.asarray(v, np.float32) for v in (x, y, z)))
    xyz = torch.from_numpy(np.ascontiguousarray(np.stack(b, -1).reshape(-1, 3))).cuda()
    out = torch.empty(xyz.shape[0], dtype=torch.float32, device="cuda")
    sh, nsh, sf, nsf = tables.arrays()
    call("ct2_phantom_points", C.cast(sh, C.c_void_p), nsh, C.cast(sf, C.c_void_p), nsf,
         _dev.ptr(xyz), _dev.ptr(out), xyz.shape[0], 0, _dev.stream_handle())
    res = out.cpu().numpy().reshape(b[0].shape)
    return float(res) if res.ndim == 0 else res


def _f32_dev(x, shape):
    """In-place target: the tensor itself when it already is a contiguous
    float32 device tensor, else a device copy."""
    if isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == torch.float32 and x.is_contiguous():
        return x.view(shape)
    return _dev.to_dev(x, torch.float32, shape)


def _store_back(holder, attr, t):
    cur = getattr(holder, attr)
    if isinstance(cur, torch.Tensor):
        if cur.data_ptr() != t.data_ptr():
            setattr(holder, attr, t.reshape(cur.shape))
    else:
        setattr(holder, attr, t.cpu().numpy().reshape(np.shape(cur)))


def _project_sino(tables, sino, views=None):
    ns, nu, nv = sino.ns, sino.nu, sino.nv
    mat = _f32_dev(sino.mat, (ns, nu, nv))
    ind = _dev.ind_dev(sino.indicator, (ns, nu, nv))
    fv = _dev.to_dev(sino.frame_vectors, torch.float32, (ns, 18))
    if views is None:
        project_views(tables, mat, ind, fv, ns, nu, nv, sino.du, sino.dv, sino.u0, sino.v0, 1)
    else:
        p = int(views)
        project_views(tables, mat[p:p + 1], None if ind is None else ind[p:p + 1], fv[p:p + 1],
                      1, nu, nv, sino.du, sino.dv, sino.u0, sino.v0, 1)
    _store_back(sino, "mat", mat)


def _project_lino(tables, lino):
    mat = _f32_dev(lino.mat, (lino.nr,))
    ind = _dev.ind_dev(lino.indicator, (lino.nr,))
    rv = _dev.to_dev(lino.ray_vectors, torch.float32, (lino.nr, 6))
    project_rays(tables, mat, ind, rv, lino.nr, 1)
    _store_back(lino, "mat", mat)


def _embed_image(tables, image):
    mat = _f32_dev(image.mat, (image.nx, image.ny, image.nz))
    embed(tables, mat, image.nx, image.ny, image.nz, image.dx, image.dy, image.dz,
          image.x0, image.y0, image.z0, 1)
    _store_back(image, "mat", mat)


# ----------------------------------------------------------- phantom model ----
class shape3D:
    """Geometric shape for building 3D phantoms (tomo3D.py:1675-1701)."""

    def __init__(self, x0=0., y0=0., z0=0., att=1.0):
        self.x0, self.y0, self.z0, self.att = x0, y0, z0, att

    def __str__(self):
        return (f"Shape parms:\nattenuation = {self.att:f}\n"
                f"center = ({self.x0:f},{self.y0:f},{self.z0:f})\n")

    def _add_to(self, tables):
        raise NotImplementedError

    def tables(self):
        t = Tables()
        self._add_to(t)
        return t

    def project_to(self, sino):
        """Adds the projection to every view of sino (tomo3D.py:2386, :2836)."""
        _project_sino(self.tables(), sino)

    def project_to_single_view(self, sino, nview):
        """tomo3D.py:2221, :2800."""
        _project_sino(self.tables(), sino, nview)

    def lino_project_to(self, lino):
        """tomo3D.py:2552, :2873."""
        _project_lino(self.tables(), lino)

    def embed_in(self, image):
        """Adds the voxelised shape to image.mat (tomo3D.py:1967, :2732)."""
        _embed_image(self.tables(), image)

    def _voxval_tables(self):
        """The tables of the shape's point test voxvalf."""
        return self.tables()

    def voxvalf(self, x, y, z):
        """The shape's value at (x, y, z) by the reference's point test
        (ellipsoid: tomo3D.py:2725-2730; general object: tomo3D.py:2122-2220);
        arrays of points are evaluated in one launch."""
        return voxval_at(self._voxval_tables(), x, y, z)


class ellipsoid(shape3D):
    """tomo3D.py:2704-2894: semi-axes ax, ay, az; rotated by beta about z, then
    alpha about the new y."""

    def __init__(self, x0=0., y0=0., z0=0., att=1.0, ax=1.0, ay=1.0, az=1.0, alpha=0.0,
                 beta=0.0):
        self.ax, self.ay, self.az = ax, ay, az
        self.alpha, self.beta = alpha, beta
        shape3D.__init__(self, x0, y0, z0, att)

    def __str__(self):
        return ("An ellipsoid!\n" + shape3D.__str__(self) +
                f"ellipsoid parms:\nax = {self.ax:f}  \nay = {self.ay:f}\naz = {self.az:f}\n"
                f"alpha = {self.alpha:f}\nbeta = {self.beta:f}\n")

    def _add_to(self, tables):
        tables.add_ellipsoid(self.att, self.ax, self.ay, self.az, self.x0, self.y0, self.z0,
                             self.alpha, self.beta)

    def voxval(self, x, y, z):
        """Inside test at a point (tomo3D.py:2765-2796; closed boundary)."""
        ca, sa, cb, sb = cos(self.alpha), sin(self.alpha), cos(self.beta), sin(self.beta)
        rx, ry, rz = x - self.x0, y - self.y0, z - self.z0
        r1x, r1y = cb * rx + sb * ry, -sb * rx + cb * ry
        r2x, r2z = ca * r1x + sa * rz, -sa * r1x + ca * rz
        lhs = (r2x / self.ax) ** 2 + (r1y / self.ay) ** 2 + (r2z / self.az) ** 2
        return self.att if lhs <= 1.0 else 0.0


class surface:
    """A bounding surface of a gen_object (tomo3D.py:1704-1794)."""
    label = "surface"

    def __init__(self, xc=0., yc=0., zc=0., alpha=0., beta=0., gamma=0.):
        self.xc, self.yc, self.zc = xc, yc, zc
        self.alpha, self.beta, self.gamma = alpha, beta, gamma
        self.xu1, self.xu2, self.xu3 = _rotate((1., 0., 0.), alpha, beta, gamma)
        self.yu1, self.yu2, self.yu3 = _rotate((0., 1., 0.), alpha, beta, gamma)
        self.zu1, self.zu2, self.zu3 = _rotate((0., 0., 1.), alpha, beta, gamma)

    def __str__(self):
        return (f"surface: {self.label} \n\n  at ({self.xc:.2f},{self.yc:.2f},{self.zc:.2f}), "
                f"unit normal ({self.zu1:.2f},{self.zu2:.2f},{self.zu3:.2f}), and x-axis along "
                f"({self.xu1:.2f},{self.xu2:.2f},{self.xu3:.2f})")

    def axes(self):
        return (0.0, 0.0, 0.0)


class planar(surface):
    """Half-space z >= 0 of the surface frame (tomo3D.py:1796-1802)."""
    label = "plane"


class ellipsoidal(surface):
    label = "ellipsoid"

    def __init__(self, xc=0., yc=0., zc=0., alpha=0., beta=0., gamma=0., ax=1., ay=1., az=1.):
        self.ax, self.ay, self.az = ax, ay, az
        surface.__init__(self, xc, yc, zc, alpha, beta, gamma)

    def axes(self):
        return (self.ax, self.ay, self.az)


class cylindrical(surface):
    label = "cylinder"

    def __init__(self, xc=0., yc=0., zc=0., alpha=0., beta=0., gamma=0., ax=1., ay=1.):
        self.ax, self.ay = ax, ay
        surface.__init__(self, xc, yc, zc, alpha, beta, gamma)

    def axes(self):
        return (self.ax, self.ay, 0.0)


class conical(surface):
    label = "cone"

    def __init__(self, xc=0., yc=0., zc=0., alpha=0., beta=0., gamma=0., ax=1., ay=1.):
        self.ax, self.ay = ax, ay
        surface.__init__(self, xc, yc, zc, alpha, beta, gamma)

    def axes(self):
        return (self.ax, self.ay, 0.0)


class gen_object(shape3D):
    """Intersection of surfaces (tomo3D.py:1835-2701)."""

    def __init__(self, x0=0., y0=0., z0=0., att=1.0, nsurf=0, surfaces=None):
        self.surfaces = list(surfaces) if surfaces is not None else []
        self.nsurf = len(self.surfaces)
        shape3D.__init__(self, x0, y0, z0, att)

    def __str__(self):
        s = f"Object center at ({self.x0:.2f},{self.y0:.2f},{self.z0:.2f})\nObject surfaces:\n"
        return s + "".join("\t" + str(i) + "\n" for i in self.surfaces)

    def add_surface(self, surface=None):
        self.surfaces.append(planar() if surface is None else surface)
        self.nsurf += 1

    def rotate(self, alpha=0., beta=0., gamma=0.):
        """Rotates every surface's centre and frame about the object centre
        (tomo3D.py:1865-1963)."""
        for s in self.surfaces:
            s.xc, s.yc, s.zc = _rotate((s.xc, s.yc, s.zc), alpha, beta, gamma)
            s.xu1, s.xu2, s.xu3 = _rotate((s.xu1, s.xu2, s.xu3), alpha, beta, gamma)
            s.yu1, s.yu2, s.yu3 = _rotate((s.yu1, s.yu2, s.yu3), alpha, beta, gamma)
            s.zu1, s.zu2, s.zu3 = _rotate((s.zu1, s.zu2, s.zu3), alpha, beta, gamma)

    def _add_to(self, tables):
        rows = []
        for kind in (CT2_SURF_PLANE, CT2_SURF_ELLIPSOID, CT2_SURF_CYLINDER, CT2_SURF_CONE):
            for s in self.surfaces:
                if _KIND_OF_LABEL.get(s.label) != kind:
                    continue
                rows.append((kind, s.axes(), (s.xu1, s.xu2, s.xu3), (s.yu1, s.yu2, s.yu3),
                             (s.zu1, s.zu2, s.zu3),
                             (s.xc + self.x0, s.yc + self.y0, s.zc + self.z0)))
        tables.add_general(self.att, rows)

    def _voxval_tables(self):
        """gen_object.voxvalf's arguments (tomo3D.py:2122-2206): each surface's
        own angles, not the frame that rotate() turned, and its centre offset
        by the object's, planes first, then ellipsoids, cylinders, cones."""
        rows = []
        for kind in (CT2_SURF_PLANE, CT2_SURF_ELLIPSOID, CT2_SURF_CYLINDER, CT2_SURF_CONE):
            for s in self.surfaces:
                if _KIND_OF_LABEL.get(s.label) != kind:
                    continue
                rows.append((kind | CT2_SURF_ANGLES, s.axes(), (s.alpha, s.beta, s.gamma),
                             (0., 0., 0.), (0., 0., 0.),
                             (s.xc + self.x0, s.yc + self.y0, s.zc + self.z0)))
        t = Tables()
        t.add_general(self.att, rows)
        return t


class phantom3D:
    """A list of additive shapes (tomo3D.py:1507-1577)."""

    def __init__(self):
        self.num_components = 0
        self.components = []

    def __str__(self):
        return "".join(f"component {i + 1}: {c}\n" for i, c in enumerate(self.components))

    def __add__(self, other):
        new = phantom3D()
        for c in self.components + other.components:
            new.add_component(c)
        return new

    def add_component(self, component):
        self.num_components += 1
        self.components.append(component)

    def tables(self):
        t = Tables()
        for c in self.components:
            c._add_to(t)
        return t

    def project_to(self, sino):
        """All components over all views in one launch (tomo3D.py:1562-1567)."""
        _project_sino(self.tables(), sino)

    def lino_project_to(self, lino):
        _project_lino(self.tables(), lino)

    def embed_in(self, image):
        _embed_image(self.tables(), image)
