"""The C-ABI library loads without a GPU and exports every entry point that
include/ct2.h declares; the ctypes prototypes cover exactly that set."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "ct2.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|size_t|const char\*)\s+(ct2_\w+)\s*\(", text)))


def test_header_parses():
    names = header_functions()
    assert "ct2_sinoproj_linear" in names and len(names) >= 30


def test_library_exports_header():
    import ctypes
    from ct2 import _lib
    if not os.path.exists(_lib.LIBPATH):
        pytest.fail("libct2.so is not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(_lib.LIBPATH)
    missing = [n for n in header_functions() if getattr(lib, n, None) is None]
    assert not missing, f"not exported: {missing}"


def header_prototypes():
    """name -> signature string in the ctypes table's letters, parsed from the header."""
    text = open(os.path.join(ROOT, "include", "ct2.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    protos = {}
    for m in re.finditer(r"\b(?:int|size_t|const char\*)\s+(ct2_\w+)\s*\(([^)]*)\)\s*;", text):
        args = [a.strip() for a in m.group(2).split(",") if a.strip() and a.strip() != "void"]
        sig = ""
        for a in args:
            if "*" in a:
                sig += "P"
            elif a.startswith("ct2_stream_t"):
                sig += "S"
            else:
                ty = a.rsplit(None, 1)[0]
                sig += {"int": "i", "float": "f", "double": "d", "size_t": "z",
                        "unsigned": "u", "unsigned int": "u"}[ty]
        protos[m.group(1)] = sig
    return protos


def test_prototypes_match_header():
    from ct2 import _lib
    assert sorted(_lib.SIGNATURES) == header_functions()
    protos = header_prototypes()
    wrong = {n: (s, protos[n]) for n, s in _lib.SIGNATURES.items() if protos[n] != s}
    assert not wrong, f"ctypes table disagrees with the header: {wrong}"


def test_error_path_without_device():
    """Invalid arguments are rejected on the host before any HIP call."""
    import ctypes as C
    from ct2 import _lib
    with pytest.raises(_lib.CT2Error, match="must be positive"):
        _lib.call("ct2_sinoproj_linear", None, None, None, 0, 4, 4, 1.0, 1.0, 0.0, 0.0,
                  None, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 4, 4, 4, None)
    with pytest.raises(_lib.CT2Error, match="volume dimension"):
        _lib.call("ct2_sinoproj_linear", None, None, None, 2, 4, 4, 1.0, 1.0, 0.0, 0.0,
                  None, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0, 4, 4, None)
    d = _lib.ct2_desc()
    _lib.call("ct2_desc_init", C.byref(d), 4, 8, 8, 1.0, 1.0, -4.0, -4.0, 8, 8, 8,
              1.0, 1.0, 1.0, -4.0, -4.0, -4.0)
    assert d.arith == _lib.CT2_ARITH_EXACT and d.interp == _lib.CT2_LINEAR
    assert _lib.lib().ct2_version() >= 100


def test_argument_errors_without_gpu():
    """Invalid sizes and null pointers are rejected with CT2_EINVAL and a
    message before any HIP call, so this runs without a device; the Python
    binding raises CT2Error carrying that message."""
    import ctypes as C
    from ct2 import _lib
    from ct2._lib import CT2Error, call, ct2_desc, ct2_stencil
    EINVAL = -1
    L = _lib.lib()

    def err(name, *args):
        rc = getattr(L, name)(*args)
        return rc, L.ct2_last_error().decode()

    rc, msg = err("ct2_sinoave", None, 0, 4, 4, None, 1, 1, 1, 1, 1, 1, 1, 1, 1, None)
    assert rc == EINVAL and "positive" in msg
    rc, msg = err("ct2_sinodist", None, 8, 8, 8, None, 3, 3, 3, 3, 3, 3, 3, 3, 3, None)
    assert rc == EINVAL and "outside" in msg          # last window ends at 9 > 8
    rc, msg = err("ct2_sinoscale", None, None, 0, 4, 4, 1.0, 1.0, 0.0, 0.0, None)
    assert rc == EINVAL and "shape" in msg
    rc, msg = err("ct2_phantom_project", None, 0, None, 0, None, None, None, 0, 4, 4,
                  1.0, 1.0, 0.0, 0.0, 0, None)
    assert rc == EINVAL and "shape" in msg
    rc, msg = err("ct2_phantom_embed", None, 0, None, 0, None, 4, 4, 4, 1.0, 1.0, 1.0,
                  0.0, 0.0, 0.0, 0, None)
    assert rc == EINVAL and "null" in msg
    rc, msg = err("ct2_total_var_grad3d", 0.0, None, None, 0, 4, 4, 1, 1e-9, None, None)
    assert rc == EINVAL and "positive" in msg
    rc, msg = err("ct2_batch_sart_dev", None, None, 1, None)
    assert rc == EINVAL and "null relaxation" in msg
    d = ct2_desc()
    call("ct2_desc_init", C.byref(d), 4, 8, 8, 1.0, 1.0, -4.0, -4.0, 8, 8, 8,
         1.0, 1.0, 1.0, -4.0, -4.0, -4.0)
    st = ct2_stencil(1, 0, 1, 1, 1, 1)                 # wu = 0
    with pytest.raises(CT2Error, match="stencil extents"):
        call("ct2_stencil_project", C.byref(d), C.byref(st), None, None, None, None, None, None)
    st = ct2_stencil(1, 2, 1, 1, 1, 1)
    with pytest.raises(CT2Error, match="null pointer"):
        call("ct2_stencil_art", C.byref(d), C.byref(st), None, 1.0, None, None, None, None, None,
             None)


def test_angle_frames_only_for_voxel_values():
    """A surface with the angle-frame flag is rejected by the projection
    entries on the host, with a message, before any HIP call."""
    import ctypes as C
    from ct2 import _lib, quadric
    t = quadric.Tables()
    t.add_general(1.0, [(quadric.CT2_SURF_PLANE | quadric.CT2_SURF_ANGLES, (0., 0., 0.),
                         (0.1, 0.2, 0.3), (0., 0., 0.), (0., 0., 0.), (0., 0., 0.))])
    sh, nsh, sf, nsf = t.arrays()
    buf = (C.c_float * 64)()
    L = _lib.lib()
    rc = L.ct2_phantom_project(C.cast(sh, C.c_void_p), nsh, C.cast(sf, C.c_void_p), nsf,
                               C.cast(buf, C.c_void_p), None, C.cast(buf, C.c_void_p), 1, 2, 2,
                               1.0, 1.0, 0.0, 0.0, 0, None)
    assert rc == -1 and "angle frames" in L.ct2_last_error().decode()
