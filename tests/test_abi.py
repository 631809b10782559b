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


def test_struct_sizes_match_c_layout():
    # sat_info: 2 ints + 2 doubles; sat_event_desc: 2 ints, 2 doubles, 2 ints
    assert ctypes.sizeof(_lib.sat_info) == 24
    assert ctypes.sizeof(_lib.sat_event_desc) == 32
    assert ctypes.sizeof(_lib.sat_grid_desc) == 8 + 4 + 4 + 6 * 8


def test_missing_library_fails_loudly(tmp_path):
    import pytest
    saved = _lib._LIB
    try:
        _lib._LIB = None
        with pytest.raises(RuntimeError, match="no CPU fallback"):
            _lib.load(str(tmp_path / "absent.so"))
    finally:
        _lib._LIB = saved
