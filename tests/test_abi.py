"""The C-ABI library loads and exports every entry point its headers declare
(no GPU needed: nothing is called)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|int64_t|double|void|char|float|TVars|const char)\s*\**\s*(\w+)\s*\(",
                       text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    from vortex_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libvortex_amd.so is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(_lib.LIB_PATH)
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        names = declared(h)
        assert names, h
        missing = [n for n in names if not hasattr(L, n)]
        assert not missing, (h, missing)


def test_binding_covers_exports():
    from vortex_amd import _lib
    names = declared("vortex.h")
    assert set(names) == set(_lib.EXPORTS), set(names) ^ set(_lib.EXPORTS)


def test_struct_layouts_match_header():
    from vortex_amd import _lib, geometry
    assert ctypes.sizeof(_lib.Params) == 20 * 8 + 3 * 8 + 2 * 4 + 2 * 8 + 2 * 4
    assert geometry.PANEL_DTYPE.itemsize == 120
