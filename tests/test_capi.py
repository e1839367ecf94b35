"""libdisco_hip.so loads on a machine without a GPU and exports every entry
point that include/disco_hip.h declares (no compute calls are made here)."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from disco_amd import capi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "disco_hip.h")


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


def _c_layout(struct, fields):
    """sizeof and offsetof of the header's struct, as gcc lays it out."""
    src = ["#include <stddef.h>", "#include <stdio.h>", '#include "disco_hip.h"',
           "int main(void) {", f'printf("%zu\\n", sizeof({struct}));']
    src += [f'printf("%zu\\n", offsetof({struct}, {f}));' for f in fields]
    src += ["return 0; }"]
    d = tempfile.mkdtemp(prefix="layout_")
    c, exe = os.path.join(d, "l.c"), os.path.join(d, "l")
    with open(c, "w") as f:
        f.write("\n".join(src))
    subprocess.run(["gcc", f"-I{os.path.join(REPO, 'include')}", c, "-o", exe], check=True)
    out = [int(v) for v in subprocess.run([exe], capture_output=True, text=True,
                                          check=True).stdout.split()]
    return out[0], out[1:]


@pytest.mark.parametrize("pystruct,cstruct", [(capi.DhParams, "dh_params"),
                                              (capi.DhGrid, "dh_grid")])
def test_struct_layout_matches_header(pystruct, cstruct):
    fields = [f for f, _ in pystruct._fields_ if not f.endswith("_")]
    size, offs = _c_layout(cstruct, fields)
    assert ctypes.sizeof(pystruct) == size
    for f, off in zip(fields, offs):
        assert getattr(pystruct, f).offset == off, f
