"""Ensembles on the GPU (dh_ens_*): four disk-planet runs of different
resolution and planet mass stepped together, every kernel of their step one
launch for all four, each with its own CFL dt, planets and diagnostics.

* each member equals the same run stepped alone through dh_timestep, bit for
  bit (the grouped launch runs the same kernel bodies on the same data);
* each member matches the reference build (oracle/_ref) of its own in.par to
  the parity bar, with its Analysis sums.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from disco_amd import configs, state_io
from helpers import PARITY_TOL, compare_states

pytestmark = pytest.mark.gpu

MEMBERS = [(64, 1e-3), (128, 1e-5), (32, 3e-4), (96, 1e-4)]
STEPS = 12


def _par(nr, q):
    return configs.get("disk_planet").par_dict(Num_R=nr, Mass_Ratio=q, Autotune=0)


@pytest.mark.parametrize("graphs", [True, False])
def test_ensemble_members_equal_their_own_runs(graphs):
    from disco_amd.host_lib import Ensemble, HostRun
    d = tempfile.mkdtemp(prefix="ens_")
    print(f"case ensemble of {len(MEMBERS)} disk-planet runs, graphs {graphs}: {STEPS} steps",
          flush=True)
    runs = [HostRun("disk_planet", state_io.write_par(os.path.join(d, f"e{i}"), _par(nr, q)))
            for i, (nr, q) in enumerate(MEMBERS)]
    solo = [HostRun("disk_planet", state_io.write_par(os.path.join(d, f"s{i}"), _par(nr, q)))
            for i, (nr, q) in enumerate(MEMBERS)]
    ens = Ensemble(runs, graphs=graphs)
    try:
        for n in range(STEPS):
            ens.step()
            for i, r in enumerate(solo):
                dt = r.getmindt()
                assert dt == ens.dts[i], (n, i, dt, ens.dts[i])
                r.timestep(dt)
        ens.sync()
        for i, (a, b) in enumerate(zip(runs, solo)):
            pa, pb = os.path.join(d, f"a{i}.dstate"), os.path.join(d, f"b{i}.dstate")
            a.write(pa)
            b.write(pb)
            x, y = state_io.read_state(pa), state_io.read_state(pb)
            for key in ("prim", "piph", "dphi", "wiph", "planets"):
                assert np.array_equal(getattr(x, key), getattr(y, key)), (i, key)
            # conserved variables of real cells (a ghost cell's are never
            # written or read: only its primitives are boundary data)
            m = x.real_mask()
            assert np.array_equal(x.cons[:, m], y.cons[:, m]), (i, "cons")
            assert x.t == y.t
            print(f"  member {i} (Nr {MEMBERS[i][0]}): identical to its own run", flush=True)
    finally:
        ens.close()
        for r in runs + solo:
            r.close()


def test_ensemble_members_match_the_reference():
    from disco_amd.host_lib import Ensemble, HostRun
    cfg = configs.get("disk_planet")
    ref = configs.ref_binary(cfg)
    if not os.path.exists(ref):
        pytest.skip(f"{ref} not built")
    d = tempfile.mkdtemp(prefix="ensref_")
    dirs = []
    for i, (nr, q) in enumerate(MEMBERS):
        di = os.path.join(d, str(i))
        state_io.write_par(di, _par(nr, q))
        subprocess.run([ref, di, str(STEPS)], cwd=di, check=True, capture_output=True)
        dirs.append(di)
    runs = [HostRun("disk_planet", os.path.join(di, "in.par")) for di in dirs]
    ens = Ensemble(runs)
    try:
        for _ in range(STEPS):
            ens.step()
        ens.sync()
        for i, (r, di) in enumerate(zip(runs, dirs)):
            r.write(os.path.join(di, "got.dstate"))
            got = state_io.read_state(os.path.join(di, "got.dstate"))
            want = state_io.read_state(os.path.join(di, "state_final.bin"))
            worst = compare_states(got, want)
            assert abs(got.t - want.t) <= 1e-12 * want.t
            _, q_want = state_io.read_diag(os.path.join(di, "diag.bin"))
            q_got = r.diagnostics(q_want.size).reshape(q_want.shape)
            scale = np.abs(q_want).sum(axis=0)
            scale = np.maximum(scale, 1e-6 * scale.max())
            err = (np.abs(q_got - q_want).sum(axis=0) / scale).max()
            print(f"  member {i} (Nr {MEMBERS[i][0]}): worst rel L1 vs reference {worst:.2e}, "
                  f"diagnostics {err:.2e}", flush=True)
            assert worst <= PARITY_TOL and err <= PARITY_TOL
    finally:
        ens.close()
        for r in runs:
            r.close()
