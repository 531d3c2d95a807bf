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


def test_struct_layout_matches_header(tmp_path):
    """The ctypes mirrors of tcad_device_desc / tcad_static_fields have the
    C layout of include/tcad_hip.h (size and every field offset), checked
    with the host C compiler."""
    import ctypes
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        import pytest
        pytest.skip("no host C compiler")
    fields = [f for f, _ in _lib.DeviceDesc._fields_]
    sfields = ["d_C", "d_nie_s", "d_nie", "d_tau_n", "d_tau_p", "d_dV", "d_an", "d_ap",
               "d_dlnnie"]
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "tcad_hip.h"',
             'int main(void) {', 'printf("%zu\\n", sizeof(tcad_device_desc));']
    lines += [f'printf("%zu\\n", offsetof(tcad_device_desc, {f}));' for f in fields]
    lines += ['printf("%zu\\n", sizeof(tcad_static_fields));']
    lines += [f'printf("%zu\\n", offsetof(tcad_static_fields, {f}));' for f in sfields]
    lines += ['return 0; }']
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run([cc, "-I", inc, str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    want = [ctypes.sizeof(_lib.DeviceDesc)]
    want += [getattr(_lib.DeviceDesc, f).offset for f in fields]
    want += [ctypes.sizeof(_lib.StaticFields)]
    want += [getattr(_lib.StaticFields, f).offset for f in sfields]
    assert got == want


def test_header_constants_match_binding():
    """Every integer constant the header defines has the same value in the
    ctypes binding (TCAD_E_* and TCAD_OK under their own names, the others
    without the TCAD_ prefix)."""
    src = open(HEADER).read()
    defs = dict(re.findall(r"^#define\s+(TCAD_\w+)\s+\(?(-?(?:0x)?[0-9a-fA-F]+)\)?", src, re.M))
    assert defs, "no constants parsed"
    for name, text in defs.items():
        value = int(text, 0)
        py = name if name == "TCAD_OK" or name.startswith("TCAD_E_") else name[len("TCAD_"):]
        assert hasattr(_lib, py), f"{name} has no binding constant {py}"
        assert getattr(_lib, py) == value, (name, getattr(_lib, py), value)
