"""ctypes binding of libct2.so (include/ct2.h).

The library is loaded from this package directory; a missing library is an
error, never a silent fallback.  Every function gets an explicit prototype
from the signature table below (P pointer, i int, f float, S stream), and
every call goes through ``call``, which raises on a negative status with the
library's own error text.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libct2.so")

CT2_LINEAR, CT2_NEAREST = 0, 1
CT2_AXIS_FREE, CT2_AXIS_Z = 0, 1
CT2_F32, CT2_F16, CT2_BF16, CT2_F16Z2 = 0, 1, 2, 3
CT2_F32_ZPLANES, CT2_F16_ZPLANES, CT2_BF16_ZPLANES = 4, 5, 6
CT2_ARITH_EXACT, CT2_ARITH_FAST = 0, 1
CT2_ROUND_NEAREST, CT2_ROUND_STOCHASTIC = 0, 1
CT2_TV_WORK_DOUBLES = 1040

_V = "iiiffff"          # ns, nu, nv, du, dv, u0, v0
_G = "ffffffiii"        # dx, dy, dz, x0, y0, z0, nx, ny, nz

SIGNATURES = {
    "ct2_version": "",
    "ct2_calib_issue": "iiiPS",
    "ct2_last_error": "",
    "ct2_desc_init": "P" + _V + "iiiffffff",
    "ct2_project": "PPPPPS",
    "ct2_backproject": "PPPPPS",
    "ct2_backproject_ws": "PPPPPPzS",
    "ct2_backproject_work_bytes": "P",
    "ct2_sart_step": "PfPPPPPS",
    "ct2_convert_volume": "PPiziuS",
    "ct2_convert_volume_zplanes": "PPiiiiiuS",
    "ct2_batch_create": "PiPPPPPPPiiS",
    "ct2_batch_destroy": "P",
    "ct2_batch_info": "PPP",
    "ct2_batch_project": "PS",
    "ct2_batch_backproject": "PS",
    "ct2_batch_sart": "PfiS",
    "ct2_batch_sart_dev": "PPiS",
    "ct2_sinoproj_linear": "PPP" + _V + "P" + _G + "S",
    "ct2_sinoproj_nearest": "PPP" + _V + "P" + _G + "S",
    "ct2_sinobackproject_linear": "PPP" + _V + "P" + _G + "S",
    "ct2_sinobackproject_nearest": "PPP" + _V + "P" + _G + "S",
    "ct2_pd_backproj_linear": "PPP" + _V + "P" + _G + "S",
    "ct2_pd_backproj_nearest": "PPP" + _V + "P" + _G + "S",
    "ct2_artstep_linear": "fPPP" + _V + "P" + _G + "S",
    "ct2_artstep_nearest": "fPPP" + _V + "P" + _G + "S",
    "ct2_emstep_linear": "PPP" + _V + "P" + _G + "S",
    "ct2_emstep_nearest": "PPP" + _V + "P" + _G + "S",
    "ct2_em_step": "PPPPPPPS",
    "ct2_art_step": "PfPPPPS",
    "ct2_sinoprojz": "PPP" + _V + "P" + _G + "S",
    "ct2_artstepz": "fPPP" + _V + "P" + _G + "S",
    "ct2_artstepz2vol": "fPPP" + _V + "P" + _G + "P" + _G + "S",
    "ct2_emstep_linearz": "PPP" + _V + "P" + _G + "S",
    "ct2_ray_sinoproj_linear": "PPPiP" + _G + "S",
    "ct2_ray_sinoproj_nearest": "PPPiP" + _G + "S",
    "ct2_ray_sinobackproject_nearest": "PPPiP" + _G + "S",
    "ct2_ray_artstep_linear": "fPPPiP" + _G + "S",
    "ct2_ray_artstep_nearest": "fPPPiP" + _G + "S",
    "ct2_ray_emstep_nearest": "PPPiP" + _G + "S",
    "ct2_total_var3d": "fPiiifPS",
    "ct2_total_var_grad3d": "fPPiiiifPS",
    "ct2_tv_value": "fPiiifPPS",
    "ct2_tv_grad": "fPPiiiifPPS",
    "ct2_distance": "PPPzPPS",
    "ct2_gmi": "PPiiifS",
    "ct2_calc_lp": "fPiiifPPS",
    "ct2_total_var3d_p": "fPiiifPPS",
    "ct2_total_dvar3d_p": "fPiiiffffPPS",
    "ct2_lp_grad3d": "fPPiiiifPPS",
    "ct2_total_p_var_grad3d": "fPPiiiifPPS",
    "ct2_total_p_dvar_grad3d": "fPPiiifffifPPS",
    "ct2_phantom_project": "PiPiPPPiiiffffiS",
    "ct2_phantom_project_rays": "PiPiPPPiiS",
    "ct2_phantom_embed": "PiPiPiiiffffffiS",
    "ct2_phantom_points": "PiPiPPziS",
    "ct2_sinoscale": "PPiiiffffS",
    "ct2_sinoave": "PiiiP" + "i" * 9 + "S",
    "ct2_sinodist": "PiiiP" + "i" * 9 + "S",
    "ct2_stencil_project": "PPPPPPPS",
    "ct2_stencil_weights": "PPPPPPPPS",
    "ct2_stencil_art": "PPPfPPPPPS",
    "ct2_fdk_sino_weight3d": "ffiiiffffffPPS",
    "ct2_fdk_sino_filter3d": "fiiiiPPifS",
    "ct2_fdk_sino_filter3d_noalloc": "fiiiiPPifS",
    "ct2_fdk_backproj_linear3d": "iii" + "f" * 8 + "dd" + "iii" + "f" * 9 + "PPS",
}

_CT = {"P": C.c_void_p, "i": C.c_int, "f": C.c_float, "S": C.c_void_p, "d": C.c_double,
       "z": C.c_size_t, "u": C.c_uint}


class ct2_desc(C.Structure):
    _fields_ = [("ns", C.c_int), ("nu", C.c_int), ("nv", C.c_int),
                ("du", C.c_float), ("dv", C.c_float), ("u0", C.c_float), ("v0", C.c_float),
                ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
                ("x0", C.c_float), ("y0", C.c_float), ("z0", C.c_float),
                ("interp", C.c_int), ("axis", C.c_int), ("vol_dtype", C.c_int),
                ("arith", C.c_int), ("view0", C.c_int), ("nviews", C.c_int),
                ("brick", C.c_int * 3), ("vstride", C.c_int)]


class ct2_stencil(C.Structure):
    _fields_ = [("ws", C.c_int), ("wu", C.c_int), ("wv", C.c_int),
                ("nss", C.c_int), ("nsu", C.c_int), ("nsv", C.c_int)]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is not built; run `python ct-2_amd/build.py`")
        L = C.CDLL(LIBPATH)
        for name, sig in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.argtypes = [_CT[c] for c in sig]
            fn.restype = C.c_int
        L.ct2_last_error.restype = C.c_char_p
        L.ct2_last_error.argtypes = []
        L.ct2_backproject_work_bytes.restype = C.c_size_t
        _LIB = L
    return _LIB


class CT2Error(RuntimeError):
    pass


def call(name, *args):
    fn = getattr(lib(), name, None)
    if fn is None:
        raise CT2Error(f"{name} is not exported by {LIBPATH}")
    rc = fn(*args)
    if rc != 0:
        msg = lib().ct2_last_error().decode(errors="replace")
        raise CT2Error(f"{name} failed ({rc}): {msg}")
    return rc


def exported_symbols():
    """Names from the header that the loaded library exports."""
    L = lib()
    return [n for n in SIGNATURES if getattr(L, n, None) is not None]
