"""The learned cons2prim guess on the GPU: k_c2p_guess's MFMA evaluation of
the model against the host build's (the same bfloat16 rounding, summed in a
different order), and GPU runs from the learned guess and/or with the
residual stopping rule against the GPU run with the reference's cons2prim."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from disco_amd import c2p_model, capi, configs, state_io
from helpers import field_errors

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpu_engine"))
import build as cpu_build  # noqa: E402

pytestmark = pytest.mark.gpu

PAR = dict(Num_R=48, Num_Z=12, R_Min=8.0, R_Max=20.0, Riemann_Solver=2, CFL=0.2)


def _initial():
    cfg = configs.get("gr_disk")
    par = cfg.par_dict(**PAR)
    exe = configs.ref_binary(cfg)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    d = tempfile.mkdtemp(prefix="gc2p_")
    state_io.write_par(d, par)
    subprocess.run([exe, d, "0"], cwd=d, check=True, capture_output=True)
    return cfg, par, state_io.read_state(os.path.join(d, "state_0000.bin"))


@pytest.fixture(scope="module")
def model():
    """A model fitted on the host build to three steps of the disk."""
    host = capi.load_library(cpu_build.build())
    cfg, par, a = _initial()
    g = capi.DiscoHip(cfg, par, a, library=host)
    xs, ys = [], []
    for _ in range(3):
        g.timestep(g.getmindt())
        f, t = g.c2p_harvest()
        k = c2p_model.physical(f, t)
        xs.append(f[k])
        ys.append(t[k])
    g.close()
    return c2p_model.fit(np.concatenate(xs), np.concatenate(ys), epochs=4, batch=256)


def test_gpu_model_matches_host_build(model):
    cfg, par, a = _initial()
    par = {**par, "C2P_Guess": 1}
    g = capi.DiscoHip(cfg, par, a)
    g.set_c2p_model(model)
    g.timestep(0.05)
    s = g.download(a)
    dev = g.c2p_predict()
    g.close()
    h = capi.DiscoHip(cfg, par, s, library=capi.load_library(cpu_build.build()))
    h.set_c2p_model(model)
    ref = h.c2p_predict()
    h.close()
    m = np.isfinite(ref).all(axis=1)
    print(f"{m.sum()} cells", flush=True)
    assert np.array_equal(np.isnan(dev[:, 0]), np.isnan(ref[:, 0]))
    dv2 = np.abs(dev[m, 0] - ref[m, 0])
    deta = np.abs(dev[m, 1] / ref[m, 1] - 1.0)
    print(f"  |d v2| median {np.median(dv2):.2e} max {dv2.max():.2e}; "
          f"|d eta/eta| median {np.median(deta):.2e} max {deta.max():.2e}", flush=True)
    assert np.median(dv2) < 1e-5 and dv2.max() < 1e-2
    assert np.median(deta) < 1e-5 and deta.max() < 1e-2


def _run(cfg, par, a, over, model=None, steps=4):
    g = capi.DiscoHip(cfg, {**par, **over}, a)
    if model is not None:
        g.set_c2p_model(model)
    stats = []
    for _ in range(steps):
        g.timestep(0.05)
        stats.append(g.c2p_stats())
    out = g.download(a)
    g.close()
    return out, stats


@pytest.mark.parametrize("guess,stop", [(1, 0), (0, 1), (1, 1)])
def test_gpu_learned_guess_and_residual_rule(model, guess, stop):
    cfg, par, a = _initial()
    print(f"case guess {guess} stop {stop}: {a.n_real()} cells, 4 steps", flush=True)
    ref, ref_stats = _run(cfg, par, a, {})
    out, stats = _run(cfg, par, a, {"C2P_Guess": guess, "C2P_Stop": stop},
                      model if guess else None)
    worst = max(field_errors(out, ref).values())
    it = [s["iterations"] / s["cells"] for s in stats]
    print(f"  worst rel L1 vs the reference rule {worst:.2e}; iterations/cell {it}; "
          f"reference {[r['iterations'] / r['cells'] for r in ref_stats]}", flush=True)
    assert worst < 1e-13
    for s in stats:
        if guess:
            assert s["learned"] > 0.9 * s["cells"], s
    assert all((x >= 11.0) if stop == 0 else (x < 6.0) for x in it)
