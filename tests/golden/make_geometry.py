"""Write tests/golden/geometry_reference.npz from the reference's own geometry
functions.  Run in the container where the reference is mounted:

    python tests/golden/make_geometry.py [/root/reference/emil_code]

It imports emil_code/configs3D.py (config_vectors_circular_conebeam :11-49,
config_vectors_circular_conebeam_flipped :196-234,
config_vectors_tomosynthesis :615-654) and emil_code/linoConfigs3D.py
(config_vectors_spherical_flatpanel :89-135), which need only numpy, and
evaluates them the way sinogram3D.__init__ does (tomo3D.py:113-147): view p
at s = s0 + p*ds with ds from get_ds (tomo3D.py:186-190), each frame stored
into a float32 array.  tomo3D.py itself does not parse under Python 3, so its
two-line ds rule is restated below.  The fixture holds inputs and outputs
only; nothing of the reference travels with it.
"""
from __future__ import annotations

import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

CASES = [  # (config, parms, ns, slen, s0, s_include_endpoints)
    ("circular_conebeam", {"radius": 500.0, "source_to_detector": 1000.0}, 90, 2 * np.pi, 0.0, 0),
    ("circular_conebeam", {"radius": 500.0, "source_to_detector": 1000.0}, 360, 2 * np.pi, 0.0, 0),
    ("circular_conebeam", {"radius": 60.0, "source_to_detector": 120.0}, 17, np.pi, 0.3, 1),
    ("circular_conebeam_flipped", {"radius": 500.0, "source_to_detector": 1000.0}, 90,
     2 * np.pi, 0.0, 0),
    ("circular_conebeam_flipped", {"radius": 5.0, "source_to_detector": 10.0}, 33, np.pi,
     -0.7, 1),
    ("tomosynthesis", {"radius": 650.0, "center_to_detector": 20.0}, 25, np.deg2rad(50),
     np.deg2rad(65), 1),
    ("tomosynthesis", {"radius": 650.0, "center_to_detector": 20.0}, 11, np.deg2rad(50),
     np.deg2rad(65), 0),
]

LINO_PARMS = {"source_radius": 300.0, "source_detector": 600.0}


def _import(path, name):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def get_ds(slen, ns, s_include_endpoints):
    """tomo3D.py:186-190."""
    return slen / (ns - 1.0) if s_include_endpoints else slen / ns


def main(refdir):
    cfg = _import(os.path.join(refdir, "configs3D.py"), "ref_configs3D")
    lino = _import(os.path.join(refdir, "linoConfigs3D.py"), "ref_linoConfigs3D")
    out = {}
    for q, (name, parms, ns, slen, s0, ends) in enumerate(CASES):
        fn = cfg.configs[name]
        ds = get_ds(slen, ns, ends)
        fv = np.zeros((ns, 18), np.float32)
        for ip in range(ns):
            fv[ip] = fn(s0 + ip * ds, parms)
        out[f"case{q}_fv"] = fv
    rng = np.random.default_rng(5)
    s = np.column_stack([rng.uniform(0.0, np.pi, 40), rng.uniform(0, 2 * np.pi, 40),
                         rng.uniform(-30, 30, 40), rng.uniform(-30, 30, 40)])
    s[0, 0] = 0.0   # the theta == 0 branch (linoConfigs3D.py:106-112)
    rv = np.array([lino.configs["ball_flatpanel"](tuple(si), LINO_PARMS) for si in s], np.float64)
    out["lino_s"] = s
    out["lino_rv"] = rv
    np.savez_compressed(os.path.join(HERE, "geometry_reference.npz"), **out)
    print(f"wrote {len(CASES)} frame-vector cases and {len(s)} ray vectors")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/emil_code")
