"""Writes the golden fixtures of tests/golden/ from the reference build.

Each case runs oracle/_ref/disco_ref_<config> (the reference timestep compiled
from its own sources by oracle/ref/Makefile) in a scratch directory on the
given in.par overrides, for `steps` steps, and stores
    init_*   the full state after setup and the initial boundaries,
    final_*  the full state after the steps,
    dt       the CFL timestep the reference took before each step,
as a compressed .npz (arrays only; no reference source is stored).

Run here, where /root/reference is mounted:
    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from disco_amd import configs, state_io  # noqa: E402

# name -> (config, in.par overrides, steps)
CASES = {
    "euler_yee_nr24": ("euler_yee", {"Num_R": 24}, 10),
    "euler_yee_hll_nr16": ("euler_yee", {"Num_R": 16, "Riemann_Solver": 0}, 6),
    "euler_yee_n1_nr16": ("euler_yee_n1", {"Num_R": 16}, 6),
    "euler_vortex_nr16": ("euler_vortex", {"Num_R": 16}, 10),
    # 3-D MHD Keplerian disk with constrained transport, 8 x 2 real annuli.
    "mhd_kep_r8z2": ("mhd_kep", {"Num_R": 8, "Num_Z": 2}, 12),
}


def run_reference(cfg, par, steps, outdir):
    state_io.write_par(outdir, par)
    exe = configs.ref_binary(cfg)
    subprocess.run([exe, outdir, str(steps)], cwd=outdir, check=True,
                   capture_output=True, text=True)
    a = state_io.read_state(os.path.join(outdir, "state_0000.bin"))
    b = state_io.read_state(os.path.join(outdir, "state_final.bin"))
    dts = np.loadtxt(os.path.join(outdir, "dt.txt"), ndmin=1)
    return a, b, dts


def pack(prefix, s):
    return {f"{prefix}_prim": s.prim, f"{prefix}_cons": s.cons, f"{prefix}_piph": s.piph,
            f"{prefix}_dphi": s.dphi, f"{prefix}_wiph": s.wiph,
            f"{prefix}_phi_flux": s.phi_flux}


def main():
    for name, (cname, over, steps) in CASES.items():
        cfg = configs.get(cname)
        par = cfg.par_dict(**over)
        with tempfile.TemporaryDirectory() as d:
            a, b, dts = run_reference(cfg, par, steps, d)
        data = dict(pack("init", a))
        data.update(pack("final", b))
        data["dt"] = dts
        data["np"] = a.np_
        data["r_jph"] = a.r_jph
        data["z_kph"] = a.z_kph
        data["ng"] = np.array(a.ng, dtype=np.int32)
        data["dims"] = np.array([a.nq, a.nf, a.nr, a.nz], dtype=np.int32)
        data["phi_max"] = np.array([a.phi_max])
        data["planets"] = a.planets
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **data)
        print(f"{name}: {a.n_real()} cells, {steps} steps -> {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
