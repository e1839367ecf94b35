"""Drop-in check: the same in.par given to the reference build (oracle/_ref)
and to this project's `disco` driver (host C layer + libdisco_hip.so, with the
same Initial/Planets plugins) gives the same run.

Both sides set the mesh up themselves (random ring offsets from srand(314159),
initial() on every cell), take their own CFL steps and stop after N steps.
The Analysis module's per-annulus sums are compared as well: each column
against its own L1 norm, floored at 1e-6 of the largest column's, since
columns such as the radial mass flux are small differences of the disk's
motion."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from disco_amd import configs, state_io
from helpers import PARITY_TOL, assert_reports_match, compare_states, read_report, rel_l1

pytestmark = pytest.mark.gpu
HOST = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                    "disco_amd", "host")


@pytest.mark.parametrize("cname,over,steps", [
    ("euler_yee", {}, 20),
    ("euler_yee_n1", {"Num_R": 64, "Riemann_Solver": 0}, 10),
    ("mhd_kep", {"Num_R": 32, "Num_Z": 8}, 6),
    ("mhd_kep_p2", {"Num_R": 32, "Num_Z": 8, "Max_DT": 0.01}, 6),
    # the memory-filling disk's setup (toroidal zero-net-flux field at beta
    # 100) against the reference build over 10 CFL steps
    ("mhd_kep", {**configs.C5_PAR, "Num_R": 32, "Num_Z": 10}, 10),
    # a disk-planet ensemble member (alpha viscosity, a planet on its orbit)
    ("disk_planet", {"Num_R": 128}, 20),
    # the same in a frame rotating about an axis off the origin (rotframe.c),
    # and a nozzle (sink.c) feeding gas and passive scalar into the vortex
    # and into the magnetised disk
    ("disk_planet", {"Num_R": 128, "RotFrame": 1, "RotOmega": 0.5, "RotD": 0.3}, 20),
    ("euler_yee_n1", {"Num_R": 64, "Nozzle_Type": 1, "Nozzle_Par1": 0.5, "Nozzle_Par2": 0.2,
                      "Nozzle_Par3": 0.5, "Nozzle_Par4": 5.0}, 10),
    ("mhd_kep", {"Num_R": 32, "Num_Z": 8, "Nozzle_Type": 1, "Nozzle_Par1": 1.2,
                 "Nozzle_Par2": 0.25, "Nozzle_Par3": 0.8, "Nozzle_Par4": 10.0}, 6),
    # GRMHD trajectories with HLL and HLLD over the steps a one-ulp libm change
    # moves by less than 1e-11, and GR HLLC on the 2-D disk falling in on a
    # static mesh, whose contacts stay away from the faces' speeds
    # (tests/test_conditioning.py); longer runs are checked step by step in
    # test_gpu_grmhd.py
    ("gr_disk", {"Num_R": 32, "Num_Z": 8, "Riemann_Solver": 0}, 3),
    ("gr_disk", {"Num_R": 32, "Num_Z": 8, "Riemann_Solver": 2}, 3),
    ("gr_disk", {"Num_R": 96, "Num_Z": 1, "Mesh_Motion": 0, "Init_Par0": 1}, 6),
    ("gr_disk3d", {"Num_R": 32, "Num_Z": 8, "Riemann_Solver": 0}, 3),
    # the Fishbone-Moncrief torus around a Kerr black hole (a = 0.9) with the
    # relativistic HLLD, and with HLL on a static mesh
    ("kerr_torus", {"Num_R": 32, "Num_Z": 16}, 6),
    ("kerr_torus", {"Num_R": 32, "Num_Z": 16, "Riemann_Solver": 0, "Mesh_Motion": 0}, 6),
])
def test_disco_driver_matches_reference(cname, over, steps):
    cfg = configs.get(cname)
    par = cfg.par_dict(**over)
    ref = configs.ref_binary(cfg)
    exe = os.path.join(HOST, f"disco_{cname}")
    if not os.path.exists(ref):
        pytest.skip(f"{ref} not built")
    d = tempfile.mkdtemp(prefix="dropin_")
    state_io.write_par(d, par)
    print(f"case {cname} {over}: {steps} steps", flush=True)
    subprocess.run([ref, d, str(steps)], cwd=d, check=True, capture_output=True)
    r = subprocess.run([exe, str(steps)], cwd=d, check=True, capture_output=True, text=True,
                       timeout=300)
    print("  " + r.stdout.strip().splitlines()[-1], flush=True)
    want = state_io.read_state(os.path.join(d, "state_final.bin"))
    got = state_io.read_state(os.path.join(d, "output.dstate"))
    assert got.step == steps
    np.testing.assert_array_equal(got.np_, want.np_)
    np.testing.assert_array_equal(got.r_jph, want.r_jph)
    worst = compare_states(got, want)
    assert abs(got.t - want.t) <= 1e-12 * want.t
    print(f"  worst rel L1 vs reference {worst:.2e}", flush=True)
    if want.nf:
        m = want.real_mask()
        ephi = rel_l1(got.phi_flux[:, m], want.phi_flux[:, m])
        print(f"  magnetic fluxes rel L1 {ephi:.2e}", flush=True)
        assert ephi <= PARITY_TOL
    t_want, q_want = state_io.read_diag(os.path.join(d, "diag.bin"))
    t_got, q_got = state_io.read_diag(os.path.join(d, "output.diag"))
    assert abs(t_got - t_want) <= 1e-12 * t_want and q_got.shape == q_want.shape
    scale = np.abs(q_want).sum(axis=0)
    scale = np.maximum(scale, 1e-6 * scale.max())
    err = np.abs(q_got - q_want).sum(axis=0) / scale
    print(f"  diagnostics ({cfg.analysis}) worst column {err.max():.2e}", flush=True)
    assert err.max() <= PARITY_TOL
    # report.c's line for the final state (report_ref.dat) against the
    # driver's last report.dat line, to the digits report.c prints
    rep = read_report(os.path.join(d, "report.dat"))
    assert rep[0, 0] == 0.0
    assert_reports_match(rep[-1], read_report(os.path.join(d, "report_ref.dat"))[-1])
