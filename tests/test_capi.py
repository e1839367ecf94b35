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


def _defined_dynamic_symbols(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.strip()}


def test_host_library_symbols_do_not_collide_with_kept_reference_objects():
    """A Disco binding keeps the reference's domain.o and report.o (check_dt,
    report, possiblyOutput) and leaves out the step's objects and
    analysis.o, whose functions the host layer supplies: the host library
    defines the step's reference entry points and none of the kept
    objects' symbols."""
    host = os.path.join(REPO, "disco_amd", "host", "libdisco_host_mhd_kep.so")
    if not os.path.exists(host):
        pytest.fail(f"{host} missing: run python -m disco_amd.build")
    syms = _defined_dynamic_symbols(host)
    for name in ("getmindt", "timestep", "add_diagnostics", "avg_diagnostics",
                 "zero_diagnostics", "disco_setup", "disco_sync_to_host", "disco_report",
                 "disco_check_dt"):
        assert name in syms, name
    for name in ("report", "check_dt", "possiblyOutput", "setupDomain", "read_par_file"):
        assert name not in syms, name
