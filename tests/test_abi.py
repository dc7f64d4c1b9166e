"""CPU checks of the C-ABI library: it is built, loads, and exports every
symbol include/wgrid_hip.h declares (no compute calls without a GPU)."""
import ctypes
import os
import re

from wsclean_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "wgrid_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(wgx?_[a-z0-9_]+)\s*\(", src)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    # the ctypes signature table covers the same set
    assert sorted(n for n, _, _ in _lib.SIGNATURES) == names


def test_params_struct_layout():
    # the C struct: 4 size_t, 4 double, size_t, double, size_t, int (+pad)
    assert ctypes.sizeof(_lib.WgParams) == 4 * 8 + 4 * 8 + 8 + 8 + 8 + 8


def test_library_is_gfx950():
    so = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in so
