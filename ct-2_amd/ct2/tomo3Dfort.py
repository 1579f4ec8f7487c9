"""Drop-in for the voxel embedders of the reference's f2py module
``tomo3Dfort`` (emil_code/tomo3DfortranAux.f, built by emil_code/Makefile as
``tomo3Dfort``).  The volumes are computed by ``ct2_phantom_embed`` on the
current HIP device; numpy in gives numpy out, as the f2py wrappers do."""
from __future__ import annotations

import torch

from . import _dev, quadric


def _embed(t, nx, ny, nz, dx, dy, dz, x0i, y0i, z0i):
    vol = torch.empty((nx, ny, nz), dtype=torch.float32, device=_dev.device())
    quadric.embed(t, vol, nx, ny, nz, dx, dy, dz, x0i, y0i, z0i, 0)
    return vol.cpu().numpy()


def ellipsoid_embed(nx, ny, nz, dx, dy, dz, x0i, y0i, z0i, att, alpha, beta, x0, y0, z0,
                    ax, ay, az):
    """tomo3DfortranAux.f:212-267: att at voxel centres strictly inside the
    ellipsoid, 0 elsewhere; called from emil_code/tomo3D.py:2758."""
    t = quadric.Tables()
    t.add_ellipsoid(att, ax, ay, az, x0, y0, z0, alpha, beta)
    return _embed(t, nx, ny, nz, dx, dy, dz, x0i, y0i, z0i)


def gen_obj_embed(nx, ny, nz, dx, dy, dz, x0i, y0i, z0i, att, nsurf, *surf):
    """tomo3DfortranAux.f:420-582; called from emil_code/tomo3D.py:2106."""
    t = quadric.tables_from_f2py(att, nsurf, *surf)
    return _embed(t, nx, ny, nz, dx, dy, dz, x0i, y0i, z0i)
