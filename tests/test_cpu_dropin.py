"""Drop-in check on the CPU: the same in.par given to the reference build
(oracle/_ref) and to this project's `disco` driver -- the host layer with
the Initial/Planets plugins, here linked to the host build of the engine --
gives the same run bit for bit: mesh set-up with the reference's rand()
sequence, initial data and noise, the initial magnetic fluxes
(set_B_fields), CFL steps, the timesteps themselves and the Analysis
module's per-annulus sums (diag_euler with and without a companion planet,
diag_disk2d, diag_mhdmri) -- for Euler, Newtonian MHD and GRMHD.  The GPU version of
this test (test_gpu_dropin.py) runs the driver on libdisco_hip.so."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from disco_amd import configs, state_io

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpu_engine"))
import build as cpu_build  # noqa: E402


@pytest.mark.parametrize("cname,over,steps", [
    ("euler_yee", {"Num_R": 24}, 6),
    ("euler_yee_n1", {"Num_R": 16, "Riemann_Solver": 0}, 4),
    ("mhd_kep", {"Num_R": 8, "Num_Z": 4}, 5),
    ("mhd_kep", {"Num_R": 12, "Num_Z": 6, "Max_DT": 0.02}, 6),
    ("mhd_kep_p2", {"Num_R": 12, "Num_Z": 4, "Max_DT": 0.02}, 4),
    # GRMHD on the equator of a Schwarzschild black hole: HLLC, HLL, and mesh
    # motion from the fluid's angular velocity (get_omega)
    ("gr_disk", {"Num_R": 24, "Num_Z": 6}, 8),
    ("gr_disk", {"Num_R": 16, "Num_Z": 4, "Riemann_Solver": 0}, 4),
    ("gr_disk", {"Num_R": 16, "Num_Z": 4, "Mesh_Motion": 1}, 4),
    # the relativistic HLLD: a smooth run, one where the reference falls back
    # to the HLL state at some faces, and a strong field with many fallbacks
    ("gr_disk", {"Num_R": 16, "Num_Z": 4, "Riemann_Solver": 2, "CFL": 0.2}, 8),
    ("gr_disk", {"Num_R": 24, "Num_Z": 6, "Riemann_Solver": 2}, 4),
    ("gr_disk", {"Num_R": 24, "Num_Z": 6, "Riemann_Solver": 2, "Init_Par1": 10.0, "CFL": 0.3}, 6),
])
def test_driver_bit_exact_vs_reference(cname, over, steps):
    cfg = configs.get(cname)
    ref = configs.ref_binary(cfg)
    if not os.path.exists(ref):
        pytest.skip(f"{ref} not built")
    exe = cpu_build.build_driver(cname)
    par = cfg.par_dict(**over)
    d = tempfile.mkdtemp(prefix="cpudropin_")
    state_io.write_par(d, par)
    subprocess.run([ref, d, str(steps)], cwd=d, check=True, capture_output=True)
    subprocess.run([exe, str(steps)], cwd=d, check=True, capture_output=True, timeout=300)
    want = state_io.read_state(os.path.join(d, "state_final.bin"))
    got = state_io.read_state(os.path.join(d, "output.dstate"))
    assert got.step == steps and got.t == want.t
    np.testing.assert_array_equal(got.np_, want.np_)
    m = want.real_mask()
    for key in ("prim", "cons"):
        a, b = getattr(got, key)[:, m], getattr(want, key)[:, m]
        assert np.array_equal(a, b), f"{key}: max diff {np.abs(a - b).max():.3e}"
    assert np.array_equal(got.piph, want.piph)
    if want.nf:
        assert np.array_equal(got.phi_flux, want.phi_flux)
    t_want, q_want = state_io.read_diag(os.path.join(d, "diag.bin"))
    t_got, q_got = state_io.read_diag(os.path.join(d, "output.diag"))
    assert q_got.shape == q_want.shape == (got.nr * got.nz, configs.DIAG_KINDS[cfg.analysis][1])
    assert t_got == t_want and np.array_equal(q_got, q_want)
    if len(want.planets) > 1:
        assert np.any(q_want[:, 3] != 0.0)  # the torque column reads planet 1
