"""report (report.c:7-202; dh_report) beyond the reference's own report line
(compared in test_cpu_dropin.py / test_gpu_dropin.py to the digits it
prints): closed forms on uniform states, where every column is fixed by the
mesh's total volume, and the GPU's per-wavefront sums against the host
build's in-order sums on a disk with a planet."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from disco_amd import capi, configs, state_io

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpu_engine"))
import build as cpu_build  # noqa: E402

COL = {n: i for i, n in enumerate(capi.DiscoHip.REPORT_COLUMNS)}


def _initial(cname, over):
    cfg = configs.get(cname)
    par = cfg.par_dict(**over)
    exe = configs.ref_binary(cfg)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    d = tempfile.mkdtemp(prefix="report_")
    state_io.write_par(d, par)
    subprocess.run([exe, d, "0"], cwd=d, check=True, capture_output=True)
    return cfg, par, state_io.read_state(os.path.join(d, "state_0000.bin"))


def _real_extent(s):
    """Edges of the real annuli: r_jph[j + 1] is edge j (edge -1 first)."""
    nga, ngb, nza, nzb = s.ng
    r1, r2 = s.r_jph[nga], s.r_jph[s.nr - ngb]
    z1, z2 = s.z_kph[nza], s.z_kph[s.nz - nzb]
    return r1, r2, z1, z2


@pytest.mark.parametrize("cname,over,B", [
    ("euler_yee", {"Num_R": 24}, None),
    ("mhd_kep", {"Num_R": 10, "Num_Z": 4}, (0.3, -0.2, 0.5)),
])
def test_uniform_state_closed_forms(cname, over, B):
    """rho = 2, P = 3, u = 0 (and a uniform B): Mass = 2 V and Vol = V =
    (r2^2 - r1^2) phi_max (z2 - z1) / 2; S_R = the volume-weighted mean
    radius (2/3)(r2^3 - r1^3)/(r2^2 - r1^2), which the cells' centroids give
    exactly; L1_rho = 1, L1_isen = |3/2^gamma - 1|, Mdot = 0; the field
    columns |Br|, Br^2 V / 2, Br Bp / P and P / (B^2 / 2)."""
    cfg, par, s = _initial(cname, over)
    s.prim[:] = 0.0
    s.prim[0], s.prim[1] = 2.0, 3.0
    if B is not None:
        s.prim[5], s.prim[6], s.prim[7] = B
    g = capi.DiscoHip(cfg, par, s, library=capi.load_library(cpu_build.build()))
    out = g.report()
    g.close()
    r1, r2, z1, z2 = _real_extent(s)
    V = 0.5 * (r2 * r2 - r1 * r1) * s.phi_max * (z2 - z1)
    gam = float(par.get("Adiabatic_Index", 5.0 / 3.0))
    assert out[COL["Mass"]] == pytest.approx(2.0 * V, rel=1e-13)
    assert out[COL["rho_min"]] == out[COL["rhoavg_min"]] == 2.0
    assert out[COL["S_R"]] == pytest.approx(2.0 / 3.0 * (r2**3 - r1**3) / (r2**2 - r1**2),
                                            rel=1e-13)
    assert out[COL["L1_rho"]] == pytest.approx(1.0, rel=1e-13)
    assert out[COL["L1_isen"]] == pytest.approx(abs(3.0 / 2.0**gam - 1.0), rel=1e-13)
    assert out[COL["Mdot"]] == 0.0
    assert abs(out[COL["PsiR"]]) < 1e-13 * V and abs(out[COL["PsiI"]]) < 1e-13 * V
    assert out[COL["Torque"]] == out[COL["Power"]] == out[COL["Fr"]] == 0.0
    if B is None:
        assert out[COL["L1_B"]] == out[COL["Br2"]] == out[COL["aM"]] == 0.0
        assert np.isinf(out[COL["bM"]])
    else:
        br, bp, bz = B
        assert out[COL["L1_B"]] == pytest.approx(abs(br), rel=1e-13)
        assert out[COL["Br2"]] == pytest.approx(0.5 * br * br * V, rel=1e-13)
        assert out[COL["aM"]] == pytest.approx(br * bp / 3.0, rel=1e-13)
        assert out[COL["bM"]] == pytest.approx(3.0 / (0.5 * (br * br + bp * bp + bz * bz)),
                                               rel=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("cname,over", [
    ("disk_planet", {"Num_R": 128}),
    ("mhd_kep", {"Num_R": 32, "Num_Z": 8}),
])
def test_gpu_report_matches_host_build(cname, over):
    """Three steps on the GPU, then dh_report there and on the host build
    from the downloaded state: the per-wavefront and in-order sums agree to
    1e-13 of each column's scale (the total of its terms' magnitudes is at
    most the disk's mass or volume times the largest cell term)."""
    cfg, par, s = _initial(cname, over)
    print(f"case {cname} {over}: {len(s.planets)} planets", flush=True)
    g = capi.DiscoHip(cfg, par, s)
    for _ in range(3):
        g.timestep(g.getmindt())
    st = g.download(s)
    pl = st.planets if len(st.planets) else None
    dev = g.report(pl)
    g.close()
    h = capi.DiscoHip(cfg, par, st, library=capi.load_library(cpu_build.build()))
    ref = h.report(pl)
    h.close()
    if len(s.planets) > 1:
        assert ref[COL["Torque"]] != 0.0
    scale = max(abs(ref[COL["Mass"]]), 1.0)
    for name, a, b in zip(capi.DiscoHip.REPORT_COLUMNS, dev, ref):
        if not np.isfinite(b):
            assert a == b, name
            continue
        print(f"  {name:10s} {a: .15e} {b: .15e}", flush=True)
        assert abs(a - b) <= 1e-13 * max(abs(b), scale), (name, a, b)
