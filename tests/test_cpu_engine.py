"""The engine's kernel bodies, built as host code (tests/cpu_engine), against
the oracle.

The GPU library and this host build share every kernel body
(disco_amd/csrc/dh_kernels.h) and the launch sequence (dh_pipeline.h); only
the executor differs.  Compiled without FMA contraction against the host
libm, the owner-computes restatement -- merge-path faces, gathers in the
reference's summation order, the edge-EMF and magnetic-flux updates, the
merge-path edge walk and the face flips -- must reproduce the oracle bit for
bit.  On the GPU the same code differs only by FMA contraction and the
device libm (checked to 1e-10 in the gpu tests).
"""
import ctypes as C
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from disco_amd import capi, configs, state_io
from helpers import GOLDEN_CASES, div_b, load_golden
from oracle.bind import Oracle

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpu_engine"))
import build as cpu_build  # noqa: E402


@pytest.fixture(scope="module")
def cpu_lib():
    lib = capi.load_library(cpu_build.build())
    lib.dhx_flip_count.argtypes = [C.c_void_p]
    lib.dhx_flip_count.restype = C.c_long
    for f in ("dhx_bc_rings", "dhx_bc_misses"):
        getattr(lib, f).argtypes = [C.c_void_p]
        getattr(lib, f).restype = C.c_long
    return lib


def _assert_same(out, ref, label):
    m = ref.real_mask()
    for key in ("prim", "cons"):
        a, b = getattr(out, key)[:, m], getattr(ref, key)[:, m]
        bad = np.argwhere(a != b)
        assert bad.size == 0, f"{label}: {key} differs at (q, cell) {bad[:3].tolist()}"
    assert np.array_equal(out.piph, ref.piph), f"{label}: piph differs"
    assert np.array_equal(out.dphi, ref.dphi), f"{label}: dphi differs"
    if ref.nf:
        assert np.array_equal(out.phi_flux, ref.phi_flux), f"{label}: Phi differs"


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_cpu_build_bit_exact_on_fixtures(cpu_lib, name):
    cfg, par, a, b, dts = load_golden(name)
    g = capi.DiscoHip(cfg, par, a, library=cpu_lib)
    o = Oracle(cfg, par, a)
    for n, dt in enumerate(dts):
        assert g.getmindt() == dt, f"step {n}: CFL dt"
        g.timestep(dt)
        o.timestep(dt)
    out = g.download(a)
    _assert_same(out, o.get_state(), name)
    # and the reference fixture itself, on real cells
    m = b.real_mask()
    assert np.array_equal(out.prim[:, m], b.prim[:, m])
    if a.nf:
        assert cpu_lib.dhx_flip_count(g.ctx) > 0, "no face flipped: the flip path was not run"
    g.close()


def _ref_initial(cname, over):
    cfg = configs.get(cname)
    par = cfg.par_dict(**over)
    exe = configs.ref_binary(cfg)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    d = tempfile.mkdtemp(prefix="cpuref_")
    state_io.write_par(d, par)
    subprocess.run([exe, d, "0"], cwd=d, check=True, capture_output=True)
    return cfg, par, state_io.read_state(os.path.join(d, "state_0000.bin"))


@pytest.mark.parametrize("cname,over,steps", [
    ("euler_yee", {"Num_R": 16, "Num_Z": 4}, 4),              # z faces and z ghosts (held)
    # CT with flips and diode z sides.  Max_DT keeps the Keplerian shear per
    # step below one cell, as the flip hand-over (one crossing per face and
    # step) needs for div B to stay at round-off; the CFL step of a mesh
    # this coarse would shear edges across several cells per step.
    ("mhd_kep", {"Num_R": 12, "Num_Z": 6, "Max_DT": 0.02}, 20),
    ("mhd_kep", {"Num_R": 8, "Num_Z": 1}, 6),                  # CT without z faces
    # the memory-filling disk's setup (toroidal zero-net-flux field) at CFL steps
    ("mhd_kep", {**configs.C5_PAR, "Num_R": 16, "Num_Z": 6}, 10),
    ("mhd_kep", {"Num_R": 10, "Num_Z": 4, "Timestep": 1}, 4),  # forward Euler
])
def test_cpu_build_bit_exact_vs_oracle(cpu_lib, cname, over, steps):
    cfg, par, a = _ref_initial(cname, over)
    g = capi.DiscoHip(cfg, par, a, library=cpu_lib)
    o = Oracle(cfg, par, a)
    for n in range(steps):
        dt = o.getmindt()
        assert g.getmindt() == dt, f"step {n}: CFL dt"
        g.timestep(dt)
        o.timestep(dt)
    out = g.download(a)
    _assert_same(out, o.get_state(), f"{cname} {over}")
    if cfg.dh_bc(par)[2] in (configs.BC_ZEROGRAD, configs.BC_DIODE):
        # the GPU's parallel ring walk guesses every ring right (no ring
        # needs its sequential fallback)
        assert cpu_lib.dhx_bc_rings(g.ctx) > 0
        assert cpu_lib.dhx_bc_misses(g.ctx) == 0
    if a.nf and par.get("Max_DT"):
        assert cpu_lib.dhx_flip_count(g.ctx) > 0
        faces = {d: g.faces(d)[:3] for d in (1, 2) if (a.nr if d == 1 else a.nz) > 1}
        div = div_b(out, faces)[out.real_mask()]
        assert div.max() < 1e-12, f"div B {div.max():.2e}"
    g.close()


@pytest.mark.parametrize("cname,over,steps", [
    ("euler_yee", {"Num_R": 24}, 5),
    ("mhd_kep", {"Num_R": 12, "Num_Z": 6, "Max_DT": 0.02}, 5),
])
def test_face_reuse_equals_rebuild(cpu_lib, cname, over, steps):
    """An RK2 step's first substep takes the faces the previous step's
    second substep built (no substep between them moves the mesh); a run
    whose state is uploaded again before every step, which makes every step
    build its faces afresh, is bit-identical."""
    cfg, par, a = _ref_initial(cname, over)
    g = capi.DiscoHip(cfg, par, a, library=cpu_lib)
    h = capi.DiscoHip(cfg, par, a, library=cpu_lib)
    for _ in range(steps):
        dt = g.getmindt()
        assert h.getmindt() == dt
        g.timestep(dt)
        h.upload(h.download(a))
        h.timestep(dt)
    x, y = g.download(a), h.download(a)
    g.close()
    h.close()
    for key in ("prim", "cons", "piph", "dphi", "phi_flux"):
        assert np.array_equal(getattr(x, key), getattr(y, key)), key


def test_cpu_build_faces_match_oracle(cpu_lib):
    cfg, par, a = _ref_initial("mhd_kep", {"Num_R": 8, "Num_Z": 4})
    g = capi.DiscoHip(cfg, par, a, library=cpu_lib)
    o = Oracle(cfg, par, a)
    for dim in (1, 2):
        L, R, LR, cm, dA = g.faces(dim)
        f = o.faces(dim)
        assert np.array_equal(L, f["L"]) and np.array_equal(R, f["R"])
        assert np.array_equal(LR, f["LRtype"])
        assert np.array_equal(dA, f["dA"]) and np.array_equal(cm[:, 1], f["phic"])
    g.close()


def test_galloping_search_equals_bisection(cpu_lib):
    """last_true (dh_mesh.h) finds the boundary of a monotone predicate from
    any estimate: every boundary and estimate of lists up to 40 long."""
    cpu_lib.dhx_last_true.argtypes = [C.c_int, C.c_int, C.c_char_p]
    cpu_lib.dhx_last_true.restype = C.c_int
    for n in range(0, 41):
        for ans in range(0, n + 1):
            truth = bytes([1] * (ans + 1) + [0] * (n - ans))
            for g in range(-2, n + 3):
                assert cpu_lib.dhx_last_true(n, g, truth) == ans, (n, ans, g)


def test_diagnostics_accumulate_average_zero(cpu_lib):
    """add_diagnostics adds dt-weighted annulus means, avg_diagnostics divides
    by the accumulated time, zero_diagnostics clears (analysis.c:8-106)."""
    cfg, par, a = _ref_initial("mhd_kep", {"Num_R": 8, "Num_Z": 4})
    g = capi.DiscoHip(cfg, par, a, library=cpu_lib)
    g.add_diagnostics("diag_mhdmri", 0.25, a.planets)
    one = g.diagnostics()
    g.add_diagnostics("diag_mhdmri", 0.75, a.planets)
    two = g.diagnostics()
    np.testing.assert_allclose(two, 4.0 * one, rtol=1e-14, atol=0)
    g.avg_diagnostics(1.0)
    np.testing.assert_allclose(g.diagnostics(), two, rtol=1e-15, atol=0)
    # column 0 is the annulus mean density; column 15 the mean B^2
    m = a.real_mask()
    assert np.all(one[:, 0] > 0) and np.all(one[:, 15] >= 0)
    g.zero_diagnostics()
    assert not g.diagnostics().any()
    g.close()


@pytest.mark.parametrize("cname,over", [
    ("mhd_kep", {"Num_R": 16, "Num_Z": 6}),
    ("gr_disk", {"Num_R": 24, "Num_Z": 6, "R_Min": 8.0, "R_Max": 20.0, "Riemann_Solver": 2,
                 "CFL": 0.2}),
    ("euler_yee", {"Num_R": 24}),
])
def test_no_kernel_reads_memory_nothing_wrote(cname, over):
    """Device memory starts with whatever was there: the host build filling
    every allocation with 0xff bytes (NaN doubles, -1 ints) instead of zeros
    gives the same steps bit for bit, so no kernel reads an element nothing
    wrote (the edge-EMF arrays share memory with the slopes, so this also
    checks that they are dead when reused)."""
    import subprocess
    import tempfile
    cfg = configs.get(cname)
    par = cfg.par_dict(**over)
    exe = configs.ref_binary(cfg)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    d = tempfile.mkdtemp(prefix="poison_")
    state_io.write_par(d, par)
    subprocess.run([exe, d, "0"], cwd=d, check=True, capture_output=True)
    a = state_io.read_state(os.path.join(d, "state_0000.bin"))
    lib = capi.load_library(cpu_build.build())
    outs = []
    for poison in (False, True):
        if poison:
            os.environ["DH_CPU_POISON"] = "1"
        try:
            g = capi.DiscoHip(cfg, par, a, library=lib)
            for _ in range(3):
                g.timestep(g.getmindt())
            outs.append(g.download(a))
            g.close()
        finally:
            os.environ.pop("DH_CPU_POISON", None)
    for key in ("prim", "cons", "piph", "phi_flux"):
        assert np.array_equal(getattr(outs[0], key), getattr(outs[1], key), equal_nan=True), key


@pytest.mark.parametrize("order", [2, 3, 4])
def test_strip_orders_are_permutations_and_change_nothing(cpu_lib, monkeypatch, order):
    """The strip traversal orders the GPU tuner may pick visit every cell
    and face exactly once, and a run with every kernel in one of them is
    bit-identical with the oracle (owner-computes: the order of the work
    changes nothing)."""
    cpu_lib.dhx_check_orders.argtypes = [C.c_void_p]
    cpu_lib.dhx_check_orders.restype = C.c_long
    cfg, par, a = _ref_initial("mhd_kep", {"Num_R": 12, "Num_Z": 6, "Max_DT": 0.02})
    monkeypatch.setenv("DH_CPU_ORDER", str(order))
    g = capi.DiscoHip(cfg, par, a, library=cpu_lib)
    # 6 real layers and 2 ghost layers on each side: strips of 4, 8 and all
    # 10 layers, for cells, r faces, z faces and edges
    assert cpu_lib.dhx_check_orders(g.ctx) == 12
    o = Oracle(cfg, par, a)
    for n in range(4):
        dt = o.getmindt()
        assert g.getmindt() == dt, f"step {n}: CFL dt"
        g.timestep(dt)
        o.timestep(dt)
    _assert_same(g.download(a), o.get_state(), f"order {order}")
    g.close()


@pytest.mark.parametrize("cname,over", [
    ("euler_yee", {"Num_R": 24}),
    ("mhd_kep", {"Num_R": 12, "Num_Z": 6, "Max_DT": 0.02}),
])
def test_separate_launches_equal_oracle(cpu_lib, monkeypatch, cname, over):
    """Small meshes run the mesh velocity with the phi slopes and the phi
    with the r fluxes as one launch each (every test mesh here is small);
    with them kept apart, as on large meshes (DH_CPU_MERGE=0), the steps
    are bit-identical with the oracle as well."""
    cfg, par, a = _ref_initial(cname, over)
    monkeypatch.setenv("DH_CPU_MERGE", "0")
    g = capi.DiscoHip(cfg, par, a, library=cpu_lib)
    o = Oracle(cfg, par, a)
    for n in range(4):
        dt = o.getmindt()
        assert g.getmindt() == dt, f"step {n}: CFL dt"
        g.timestep(dt)
        o.timestep(dt)
    _assert_same(g.download(a), o.get_state(), f"{cname} separate launches")
    g.close()
