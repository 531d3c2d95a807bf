"""CPU-only checks of the C-ABI boundary: the library loads and exports
every entry point include/tcad_hip.h declares (no compute calls)."""
import ctypes
import os
import re

import pytest

from tcadhip import _lib

HEADER = os.path.join(os.path.dirname(__file__), "..", "include", "tcad_hip.h")


def declared():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]*?\b(tcad_\w+)\s*\(", src, re.M)))


def test_header_declares_exported_list():
    assert set(declared()) == set(_lib.exported_symbols())


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libtcadhip.so not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared():
        assert hasattr(L, name), name
    assert _lib.lib().tcad_version().decode().startswith("tcadhip")


def test_error_mapping():
    with pytest.raises(ValueError):
        _lib.check(_lib.TCAD_E_INVALID, "x")
    with pytest.raises(IndexError):
        _lib.check(_lib.TCAD_E_INDEX, "x")
    with pytest.raises(_lib.LinearSolveError):
        _lib.check(_lib.TCAD_E_NOT_CONVERGED, "x")
