"""libdisco_hip.so loads on a machine without a GPU and exports every entry
point that include/disco_hip.h declares (no compute calls are made here)."""
import ctypes
import os
import re

import pytest

from disco_amd import capi

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "include", "disco_hip.h")


def declared_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(dh_\w+)\s*\(", txt, re.M)))


def test_header_declares_the_api():
    syms = declared_symbols()
    assert "dh_timestep" in syms and "dh_getmindt" in syms
    assert sorted(capi.EXPORTED_SYMBOLS) == syms


def test_library_exports_every_declared_symbol():
    if not os.path.exists(capi.LIB_PATH):
        pytest.fail(f"{capi.LIB_PATH} missing: run python -m disco_amd.build")
    lib = ctypes.CDLL(capi.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_param_struct_layout_matches_header():
    # dh_params: 30 ints (27 named + 3 padding) then 11 + 9 doubles.
    assert ctypes.sizeof(capi.DhParams) == 30 * 4 + 20 * 8
    assert capi.DhParams.gamma_law.offset == 120
    assert ctypes.sizeof(capi.DhGrid) == 6 * 4 + 3 * 8 + 8
