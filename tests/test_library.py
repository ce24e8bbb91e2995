"""The C-ABI library builds, loads on the CPU and exports every entry point
that include/iemic_amd.h declares (no device calls)."""
import ctypes
import os
import re

from gpu_helpers import ROOT, iemic_amd


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "iemic_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(iemic_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(iemic_amd.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_signatures_cover_the_header():
    lib = iemic_amd.load()
    for n in declared_functions():
        assert getattr(lib, n).argtypes is not None, n
