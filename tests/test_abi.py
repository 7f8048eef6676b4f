"""The C-ABI library loads and exports every entry point the public header
declares (CPU only: no compute calls)."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            src = open(os.path.join(inc, fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(ofg_[a-z_0-9]+)\s*\(", src))
    return names


def test_library_exports_declared_symbols():
    from openferric_amd import _lib
    L = C.CDLL(_lib.LIB_PATH)
    declared = _declared()
    assert declared, "no declarations found"
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, missing
    assert declared <= set(_lib.EXPORTED_SYMBOLS) | declared


def test_python_binding_lists_all_symbols():
    from openferric_amd import _lib
    assert _declared() == set(_lib.EXPORTED_SYMBOLS)


def test_version_string_without_gpu():
    from openferric_amd import _lib
    assert b"gfx950" in _lib.load().ofg_version()
