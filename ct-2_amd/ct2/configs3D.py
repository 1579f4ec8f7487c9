"""Scan geometries: trajectory parameter s -> 18-float frame
(source, trajectory tangent, detector centre, e_u, e_v, e_w).

Frames are evaluated in float64 and stored as float32 by the sinogram
(emil_code/tomo3D.py:145-147); every kernel reads the float32 copy.
"""
from __future__ import annotations

import numpy as np


def config_vectors_circular_conebeam(s, parms):
    """Circular orbit in the x-y plane, detector rows along z.
    emil_code/configs3D.py:11-49."""
    R = parms["radius"]
    sdd = parms["source_to_detector"]
    c, sn = np.cos(s), np.sin(s)
    return (R * c, R * sn, 0.0,
            -R * sn, R * c, 0.0,
            (R - sdd) * c, (R - sdd) * sn, 0.0,
            -sn, c, 0.0,
            0.0, 0.0, 1.0,
            c, sn, 0.0)


def config_vectors_circular_conebeam_flipped(s, parms):
    """Circular orbit with e_u and e_v swapped (u along z).
    emil_code/configs3D.py:196-234."""
    R = parms["radius"]
    sdd = parms["source_to_detector"]
    c, sn = np.cos(s), np.sin(s)
    return (R * c, R * sn, 0.0,
            -R * sn, R * c, 0.0,
            (R - sdd) * c, (R - sdd) * sn, 0.0,
            0.0, 0.0, 1.0,
            -sn, c, 0.0,
            c, sn, 0.0)


def config_vectors_tomosynthesis(s, parms):
    """Source arc in the y-z plane over a fixed detector at z = -c2d.
    emil_code/configs3D.py:615-654."""
    R = parms["radius"]
    c2d = parms["center_to_detector"]
    c, sn = np.cos(s), np.sin(s)
    return (0.0, R * c, R * sn,
            0.0, -R * sn, R * c,
            0.0, 0.0, -c2d,
            1.0, 0.0, 0.0,
            0.0, 1.0, 0.0,
            0.0, 0.0, 1.0)


configs = {
    "circular_conebeam": config_vectors_circular_conebeam,
    "circular_conebeam_flipped": config_vectors_circular_conebeam_flipped,
    "tomosynthesis": config_vectors_tomosynthesis,
}


def frame_vectors(config_name, parms, ns, slen, s0=0.0, s_include_endpoints=1):
    """(ns,18) float32 frames at s = s0 + p*ds, ds = slen/(ns-1) when the
    endpoints are included, slen/ns otherwise (emil_code/tomo3D.py:180-196)."""
    ds = slen / (ns - 1.0) if s_include_endpoints else slen / ns
    fn = configs[config_name]
    fv = np.zeros((ns, 18), np.float32)
    for p in range(ns):
        fv[p] = fn(s0 + p * ds, parms)
    return fv


# ---- ray-list ("linogram") geometries: s -> (source, bin) 6-vector ----
def lino_spherical_flatpanel(s, parms):
    """Source on a sphere, flat panel opposite; s = (theta, phi, u, v).
    emil_code/linoConfigs3D.py:89-135."""
    srad = parms["source_radius"]
    sdd = parms["source_detector"]
    th, ph, u, v = s
    ew = np.array([np.cos(ph) * np.sin(th), np.sin(ph) * np.sin(th), np.cos(th)])
    S = srad * ew
    if th == 0.0:
        eu = np.array([0.0, 1.0, 0.0])
        ev = np.array([0.0, 1.0, 0.0])
    else:
        up = np.array([ew[1], -ew[0], 0.0])
        eu = up / np.sqrt(up[0] ** 2 + up[1] ** 2)
        ev = np.array([ew[1] * eu[2] - ew[2] * eu[1],
                       eu[0] * ew[2] - eu[2] * ew[0],
                       ew[0] * eu[1] - ew[1] * eu[0]])
    B = (srad - sdd) * ew + u * eu + v * ev
    return (*S, *B)


lino_configs = {"ball_flatpanel": lino_spherical_flatpanel}
