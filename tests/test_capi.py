"""The C-ABI library loads and exports every entry point include/fornax_gpu.h
declares (no compute call: this runs without a GPU)."""
import ctypes
import os
import re

from fornaxamr_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names += re.findall(r"^(?:int|const char \*)\s*(fx_\w+)\s*\(", text, re.M)
    return names


def test_header_lists_entry_points():
    names = declared()
    assert len(names) >= 15
    assert set(names) == set(_native.EXPORTS)


def test_library_exports_all_symbols():
    lib = ctypes.CDLL(_native.lib_path())
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
