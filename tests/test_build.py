"""The C-ABI library exists, loads and exports every declared entry point
(no GPU needed: nothing is called)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "sdl_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sdl_[a-z_]+)\s*\(", text)))


def test_header_lists_entry_points():
    names = declared_symbols()
    assert "sdl_event_run" in names and "sdl_geom_load" in names
    from sdl_amd import runtime
    assert sorted(runtime.EXPORTS) == names


def test_library_exports_every_symbol():
    path = os.path.join(ROOT, "sdl_amd", "_lib", "libsdl_amd.so")
    if not os.path.exists(path):
        pytest.skip("libsdl_amd.so not built in this tree")
    lib = ctypes.CDLL(path)
    for name in declared_symbols():
        assert hasattr(lib, name), name
