"""The device diagnostics (the Analysis module's per-annulus sums,
analysis.c:52-106) against the host build of the same bodies, which the
drop-in tests hold bit-exact to the reference: on one state, one call each.
The device adds the cells of an annulus in a wavefront tree rather than in
order, so the two differ by rounding only."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from disco_amd import capi, configs, state_io

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpu_engine"))
import build as cpu_build  # noqa: E402

pytestmark = pytest.mark.gpu

TOL = 1e-13


@pytest.mark.parametrize("cname,over,steps", [
    ("euler_yee", {}, 3),
    ("euler_yee_n1", {"Num_R": 64}, 3),
    ("mhd_kep", {"Num_R": 32, "Num_Z": 8}, 2),
    ("mhd_kep_p2", {"Num_R": 32, "Num_Z": 8}, 2),
])
def test_gpu_diagnostics_match_host_build(cname, over, steps):
    cfg = configs.get(cname)
    par = cfg.par_dict(**over)
    exe = configs.ref_binary(cfg)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    d = tempfile.mkdtemp(prefix="gdiag_")
    state_io.write_par(d, par)
    subprocess.run([exe, d, str(steps)], cwd=d, check=True, capture_output=True)
    s = state_io.read_state(os.path.join(d, "state_final.bin"))
    print(f"case {cname} {over}: {s.n_real()} cells, {cfg.analysis}", flush=True)
    out = []
    for lib in (capi.lib(), capi.load_library(cpu_build.build())):
        g = capi.DiscoHip(cfg, par, s, library=lib)
        g.add_diagnostics(cfg.analysis, 0.01, s.planets)
        out.append(g.diagnostics())
        g.close()
    gpu, host = out
    scale = np.abs(host).sum(axis=0)
    err = np.abs(gpu - host).sum(axis=0) / np.where(scale > 0, scale, 1.0)
    print(f"  worst column rel L1 {err.max():.2e}", flush=True)
    assert err.max() <= TOL, err
