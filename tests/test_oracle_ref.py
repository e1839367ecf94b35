"""The oracle against a live run of the reference build (oracle/_ref) at the
full size of the 2-D vortex workload (skipped where oracle/_ref is absent)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from disco_amd import configs, state_io
from oracle.bind import Oracle


def _ref_run(cfg, par, steps):
    exe = configs.ref_binary(cfg)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    d = tempfile.mkdtemp(prefix="dref_")
    state_io.write_par(d, par)
    subprocess.run([exe, d, str(steps)], cwd=d, check=True, capture_output=True)
    a = state_io.read_state(os.path.join(d, "state_0000.bin"))
    b = state_io.read_state(os.path.join(d, "state_final.bin"))
    dts = np.loadtxt(os.path.join(d, "dt.txt"), ndmin=1)
    return a, b, dts


@pytest.mark.parametrize("cname,over,steps", [
    ("euler_yee", {}, 4),
    ("euler_vortex", {"Num_R": 40, "Riemann_Solver": 0}, 8),
])
def test_oracle_bitwise_live(cname, over, steps, oracle_lib):
    cfg = configs.get(cname)
    par = cfg.par_dict(**over)
    a, b, dts = _ref_run(cfg, par, steps)
    o = Oracle(cfg, par, a)
    for dt in dts:
        assert o.getmindt() == dt
        o.timestep(float(dt))
    s = o.get_state()
    m = b.real_mask()
    np.testing.assert_array_equal(s.prim, b.prim)
    np.testing.assert_array_equal(s.cons[:, m], b.cons[:, m])
    np.testing.assert_array_equal(s.piph, b.piph)
