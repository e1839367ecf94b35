"""A recorded timestep (dh_enable_graphs) against eager launches: two runs of
the drop-in host layer from the same in.par, one with its RK2 step recorded
once as a HIP graph and replayed, stepped side by side with CFL steps (so dt
changes every step, and on the disk-planet run the planet moves), and the
whole state compared bit for bit after every replay.  Each comparison
downloads both states, so eager work and synchronisation sit between the
replays.  The last steps turn the graph off again and must keep agreeing."""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

from disco_amd import c2p_model, capi, configs, state_io

pytestmark = pytest.mark.gpu


def _state(run, path):
    run.write(path)
    return state_io.read_state(path)


@pytest.mark.parametrize("cname,over,steps", [
    ("euler_yee", {}, 6),
    ("euler_yee", {"Timestep": 1}, 3),  # forward Euler: never recorded
    ("mhd_kep", {"Num_R": 32, "Num_Z": 8}, 4),
    ("disk_planet", {"Num_R": 128}, 6),
    ("gr_disk", {"Num_R": 32, "Num_Z": 8, "Riemann_Solver": 2}, 3),
    ("gr_disk", {"Num_R": 32, "Num_Z": 8, "C2P_Guess": 1, "C2P_Stop": 1,
                 "C2P_Model": c2p_model.MODEL_PATH}, 3),
])
def test_recorded_timestep_matches_eager(cname, over, steps):
    from disco_amd.host_lib import HostRun
    cfg = configs.get(cname)
    print(f"case {cname} {over}: {steps} steps", flush=True)
    d = tempfile.mkdtemp(prefix="graphs_")
    # in.par Graphs: 0 eager, 1 recorded (the driver's default); no tuner
    # steps, which run eagerly
    runs = [HostRun(cname, state_io.write_par(os.path.join(d, k),
                                              cfg.par_dict(**over, Graphs=g, Autotune=0)))
            for k, g in (("eager", 0), ("graph", 1))]
    lib = capi.lib()
    try:
        for n in range(steps + 2):
            if n == steps:
                assert lib.dh_enable_graphs(C.c_void_p(runs[1].ctx), 0) == 0
            dts = [r.getmindt() for r in runs]
            assert dts[0] == dts[1], (n, dts)
            for r in runs:
                r.timestep(dts[0])
            a = _state(runs[0], os.path.join(d, "a.dstate"))
            b = _state(runs[1], os.path.join(d, "b.dstate"))
            for name in ("prim", "cons", "piph", "dphi", "wiph", "phi_flux", "planets"):
                x, y = getattr(a, name), getattr(b, name)
                assert np.array_equal(x, y), (n, name, np.max(np.abs(x - y)))
            print(f"  step {n} (dt {dts[0]:.3e}): identical", flush=True)
    finally:
        for r in runs:
            r.close()


@pytest.mark.parametrize("name", ["euler_yee_nr24", "mhd_kep_r8z2"])
def test_face_reuse_equals_rebuild_gpu(name):
    """On the GPU: an RK2 step's first substep takes the faces the previous
    step's second substep built; a run whose state is uploaded again before
    every step (an upload clears the reuse, so every step builds its faces
    afresh) stays bit-identical to it."""
    from helpers import load_golden
    cfg, par, a, _, dts = load_golden(name)
    print(f"case {name}: {len(dts)} steps, reuse against rebuild", flush=True)
    g = capi.DiscoHip(cfg, par, a)
    h = capi.DiscoHip(cfg, par, a)
    try:
        for n, dt in enumerate(dts):
            g.timestep(dt)
            h.upload(h.download(a))
            h.timestep(dt)
            x, y = g.download(a), h.download(a)
            for key in ("prim", "cons", "piph", "dphi", "phi_flux"):
                assert np.array_equal(getattr(x, key), getattr(y, key)), (n, key)
    finally:
        g.close()
        h.close()

