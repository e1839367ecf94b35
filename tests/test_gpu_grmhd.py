"""GRMHD on the GPU step by step: every step of a run, taken by the device and
by the host build of the same bodies (which matches the reference bit for
bit, tests/test_cpu_dropin.py) from one common state, agrees to round-off.

HLL and the relativistic HLLD run the 3-D disk, in the equatorial metric
(schw_ks_eq_rpz) and in the full (r, z) one (schw_ks_rpz), whose steps a
one-ulp change of the state or of libm moves by 1e-14 to 2e-13.  GR HLLC's field flux jumps where a
contact speed sits at a face speed (tests/test_conditioning.py), which in
the 3-D disk is every z face and, with the mesh moving with the gas, every
phi face: a one-ulp change moves one step there by ~1e-3, so HLLC runs the
2-D disk falling in on a static mesh, whose contacts stay away from the
faces."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from disco_amd import capi, configs, state_io
from helpers import field_errors, rel_l1

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpu_engine"))
import build as cpu_build  # noqa: E402

pytestmark = pytest.mark.gpu

STEP_TOL = 1e-12
# The relativistic HLLD's total-pressure Newton stops at |dp / p| < 1e-12
# (hlld_grmhd.c:1092-1171), so a last-bit difference in a face's states (the
# GPU contracts multiply-adds into FMAs) can end a face's solve one
# iteration earlier or later and move its flux, and a cell's field, by
# ~1e-12 (measured: 1.6e-12 in Phi at the sixth step of the 48 x 12 disk).
HLLD_STEP_TOL = 1e-11


CASES = [
    ("gr_disk", {"Num_R": 48, "Num_Z": 12, "Riemann_Solver": 0}),
    ("gr_disk", {"Num_R": 96, "Num_Z": 1, "Riemann_Solver": 1, "Mesh_Motion": 0, "Init_Par0": 1}),
    ("gr_disk", {"Num_R": 48, "Num_Z": 12, "Riemann_Solver": 2}),
    # the full (r, z) metric, on the equator and on a slab off it; the slab
    # falls onto itself and from its sixth step its HLLD solve branches on
    # round-off (tests/test_conditioning.py), so it is compared over five
    ("gr_disk3d", {"Num_R": 48, "Num_Z": 12, "Riemann_Solver": 0}),
    ("gr_disk3d", {"Num_R": 32, "Num_Z": 8, "Riemann_Solver": 2, "Z_Min": 1.0, "Z_Max": 3.0}, 5),
    # the Fishbone-Moncrief torus in the Kerr metric (a = 0.9), relativistic HLLD
    ("kerr_torus", {"Num_R": 48, "Num_Z": 24}),
]


@pytest.mark.parametrize("cname,over,nsteps", [c if len(c) == 3 else c + (6,) for c in CASES])
def test_gpu_grmhd_steps_match_host_build(cname, over, nsteps):
    cfg = configs.get(cname)
    par = cfg.par_dict(**over)
    exe = configs.ref_binary(cfg)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    d = tempfile.mkdtemp(prefix="ggr_")
    state_io.write_par(d, par)
    subprocess.run([exe, d, "0"], cwd=d, check=True, capture_output=True)
    a = state_io.read_state(os.path.join(d, "state_0000.bin"))
    print(f"case {cname} {over}: {a.n_real()} cells, {nsteps} steps from common states",
          flush=True)
    tol = HLLD_STEP_TOL if par.get("Riemann_Solver") == 2 else STEP_TOL
    host = capi.load_library(cpu_build.build())
    g = capi.DiscoHip(cfg, par, a)
    worst = 0.0
    for n in range(nsteps):
        s = g.download(a)
        outs, dts = [], []
        for lib in (capi.lib(), host):
            h = capi.DiscoHip(cfg, par, s, library=lib)
            dts.append(h.getmindt())
            h.timestep(dts[0])
            outs.append(h.download(a))
            h.close()
        assert abs(dts[1] - dts[0]) <= 1e-11 * dts[1], (n, dts)
        e = field_errors(outs[0], outs[1])
        m = a.real_mask()
        ephi = rel_l1(outs[0].phi_flux[:, m], outs[1].phi_flux[:, m])
        hydro = max(e[f"prim{q}"] for q in range(5))
        k = max(e, key=e.get)
        worst = max(worst, hydro)
        print(f"  step {n}: hydro {hydro:.2e}, worst {e[k]:.2e} ({k}), Phi {ephi:.2e}", flush=True)
        assert hydro <= tol, (n, hydro)
        assert e[k] <= tol and ephi <= tol, (n, k, e[k], ephi)
        g.timestep(dts[0])
    g.close()
