"""The C-ABI library loads and exports every symbol include/surfatt_eik.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re

from surfattpp_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "surfatt_eik.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*int\s+(sat_\w+)\s*\(", txt, flags=re.M)))


def test_header_symbols_match_binding_list():
    assert header_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_every_header_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in header_symbols():
        assert hasattr(L, name), name


def test_struct_layouts_match_the_header(tmp_path):
    # sizes and field offsets of the ctypes mirrors against the C compiler's
    import subprocess
    structs = {"sat_info": _lib.sat_info, "sat_grid_desc": _lib.sat_grid_desc,
               "sat_event_desc": _lib.sat_event_desc, "sat_batch_desc": _lib.sat_batch_desc,
               "sat_combine_desc": _lib.sat_combine_desc}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "surfatt_eik.h"', "int main(void) {"]
    for name, cls in structs.items():
        lines.append(f'  printf("{name} %zu\\n", sizeof({name}));')
        for fname, _ in cls._fields_:
            lines.append(f'  printf("{name}.{fname} %zu\\n", offsetof({name}, {fname}));')
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                 check=True).stdout.splitlines())
    for name, cls in structs.items():
        assert int(got[name]) == ctypes.sizeof(cls), name
        for fname, _ in cls._fields_:
            assert int(got[f"{name}.{fname}"]) == getattr(cls, fname).offset, f"{name}.{fname}"


def test_missing_library_fails_loudly(tmp_path):
    import pytest
    saved = _lib._LIB
    try:
        _lib._LIB = None
        with pytest.raises(RuntimeError, match="no CPU fallback"):
            _lib.load(str(tmp_path / "absent.so"))
    finally:
        _lib._LIB = saved
