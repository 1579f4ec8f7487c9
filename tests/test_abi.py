"""The C-ABI library loads without a GPU and exports every entry point that
include/ct2.h declares; the ctypes prototypes cover exactly that set."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "ct2.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|const char\*)\s+(ct2_\w+)\s*\(", text)))


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
    for m in re.finditer(r"\b(?:int|const char\*)\s+(ct2_\w+)\s*\(([^)]*)\)\s*;", text):
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
