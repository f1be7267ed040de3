"""The C-ABI library loads and exports every symbol the header declares (no GPU)."""

import os
import re

from fieldspice_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            txt = open(os.path.join(inc, fn)).read()
            names |= set(re.findall(r"^\s*(?:int|const char\*)\s+(fs_\w+)\s*\(", txt, re.M))
    return names


def test_header_declares_the_bound_entry_points():
    names = _declared()
    assert names, "no declarations found"
    assert names == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name), name


def test_last_error_is_a_string_without_gpu():
    assert isinstance(_lib.last_error(), str)
