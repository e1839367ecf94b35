"""The learned initial guess of grmhd2's cons2prim and the residual stopping
rule (disco_amd/csrc/dh_c2pnet.h, GRMHD::cons2prim_guess), on the host build
of the engine.

The guess has no counterpart in the reference, so it is pinned by what must
hold whatever the model: the harvested targets solve noble2d's two equations
for the harvested invariants; the engine's model evaluation agrees with a
float64 evaluation of the same parameters to bfloat16 rounding; a model whose
output is unphysical falls back everywhere and leaves the run bit-identical
to the reference's cons2prim; and a run from the learned guess (or stopped by
the residual rule) agrees with the reference's to round-off."""
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

PAR = dict(Num_R=24, Num_Z=6, R_Min=8.0, R_Max=20.0, Riemann_Solver=2, CFL=0.2)


def _initial(over=None):
    cfg = configs.get("gr_disk")
    par = cfg.par_dict(**{**PAR, **(over or {})})
    exe = configs.ref_binary(cfg)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    d = tempfile.mkdtemp(prefix="c2pl_")
    state_io.write_par(d, {k: v for k, v in par.items() if not k.startswith("C2P_")})
    subprocess.run([exe, d, "0"], cwd=d, check=True, capture_output=True)
    return cfg, par, state_io.read_state(os.path.join(d, "state_0000.bin"))


@pytest.fixture(scope="module")
def host():
    return capi.load_library(cpu_build.build())


@pytest.fixture(scope="module")
def trained(host):
    """A model fitted briefly to three steps of the disk (enough for guesses
    inside the physical range), and the harvested data."""
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
    X, Y = np.concatenate(xs), np.concatenate(ys)
    return c2p_model.fit(X, Y, epochs=4, batch=256, library=host)[0], X, Y


def test_harvested_targets_solve_noble2d(trained):
    """(v^2, eta) of the harvest satisfy noble2d's equations (grmhd2.c:
    1195-1230) for the harvested invariants, to float32 precision."""
    _, X, Y = trained
    X = X.astype(np.float64)
    e = np.exp(X[:, 0])
    s2, Q, psi, n = X[:, 1] * e * e, X[:, 2] * e, X[:, 3] * e ** 3, X[:, 4]
    v2 = Y[:, 0].astype(np.float64)
    eta = e * np.exp(Y[:, 1].astype(np.float64))
    w = 1.0 / np.sqrt(1.0 - v2)
    fa = (eta + Q) ** 2 * v2 - psi * (2 * eta + Q) / eta ** 2 - s2
    fb = eta - n * (1 - v2) * (eta - w) + 0.5 * Q * (1 + v2) - 0.5 * psi / eta ** 2 - e
    assert np.max(np.abs(fa) / (eta + Q) ** 2) < 1e-5
    assert np.max(np.abs(fb) / eta) < 1e-5


def test_engine_model_matches_float64_evaluation(host, trained):
    """The engine's guesses (bfloat16 weights and activations) against the
    same parameters evaluated in float64."""
    p, _, _ = trained
    cfg, par, a = _initial({"C2P_Guess": 1})
    g = capi.DiscoHip(cfg, par, a, library=host)
    g.set_c2p_model(p)
    guess = g.c2p_predict()
    f, _ = g.c2p_harvest()
    g.close()
    m = np.isfinite(f).all(axis=1)
    t = c2p_model.predict(p, f[m])
    e = np.exp(f[m, 0].astype(np.float64))
    dv2 = np.abs(guess[m, 0] - t[:, 0])
    deta = np.abs(guess[m, 1] / (e * np.exp(t[:, 1])) - 1.0)
    assert np.isnan(guess[~m]).all()
    assert np.median(dv2) < 2e-3 and np.max(dv2) < 2e-2, (np.median(dv2), np.max(dv2))
    assert np.median(deta) < 2e-3 and np.max(deta) < 2e-2, (np.median(deta), np.max(deta))


def _run(host, over, model=None, steps=4):
    cfg, par, a = _initial()
    g = capi.DiscoHip(cfg, {**par, **over}, a, library=host)
    if model is not None:
        g.set_c2p_model(model)
    per_step = []
    for _ in range(steps):
        g.timestep(0.05)
        per_step.append(g.c2p_stats())
    out = g.download(a)
    g.close()
    return out, per_step


def test_unphysical_model_falls_back_bit_for_bit(host, trained):
    """A model whose v^2 output is 2 is rejected at every cell: the run is
    the reference's cons2prim's, bit for bit."""
    p = trained[0].copy()
    base = c2p_model.N_PARAM - 2 * c2p_model.N_OUT  # out_mu, out_sd
    p[base] = 2.0
    p[base + 2] = 0.0
    ref, _ = _run(host, {})
    out, stats = _run(host, {"C2P_Guess": 1}, p)
    for s in stats:
        assert s["learned"] == 0 and s["fallbacks"] == s["cells"], s
    assert np.array_equal(out.prim, ref.prim) and np.array_equal(out.cons, ref.cons)


@pytest.mark.parametrize("guess,stop", [(1, 0), (0, 1), (1, 1)])
def test_learned_guess_and_residual_rule_match_reference_rule(host, trained, guess, stop):
    """Each step from the reference rule's run's state, with the learned
    guess and/or the residual rule, agrees with the reference rule's step to
    1e-13; the reference rule costs its 10 extra iterations per cell, the
    residual rule does not."""
    cfg, par, a = _initial()
    g = capi.DiscoHip(cfg, par, a, library=host)
    for _ in range(4):
        s = g.download(a)
        h = capi.DiscoHip(cfg, {**par, "C2P_Guess": guess, "C2P_Stop": stop}, s, library=host)
        if guess:
            h.set_c2p_model(trained[0])
        h.timestep(0.05)
        out, st = h.download(a), h.c2p_stats()
        h.close()
        g.timestep(0.05)
        ref, rst = g.download(a), g.c2p_stats()
        worst = max(field_errors(out, ref).values())
        assert worst < 1e-13, worst
        if guess:
            assert st["learned"] + st["fallbacks"] == st["cells"]
            assert st["learned"] > 0.9 * st["cells"], st
        it = st["iterations"] / st["cells"]
        assert (it >= 11.0) if stop == 0 else (it < 6.0), it
        assert rst["iterations"] / rst["cells"] >= 11.0
    g.close()


def test_better_start_takes_the_smaller_residual(host, trained):
    """C2P_Guess 2 starts each cell from the model's guess or the previous
    primitives, whichever has the smaller residual.  From a state whose
    primitives are a poor start (velocities off by 30%) it takes the model's
    guess at a tenth of the cells or more and spends no more Newton iterations than either
    start alone; from the run's own state it keeps the previous primitives
    more often than C2P_Guess 1 does; every start reaches the same
    solution."""
    cfg, par, a = _initial()
    g = capi.DiscoHip(cfg, par, a, library=host)
    for _ in range(2):
        g.timestep(0.05)
    s = g.download(a)
    g.close()
    bad = s.copy()
    bad.prim[2:5] *= 1.3
    for state, poor in ((bad, True), (s, False)):
        outs, stats = {}, {}
        for mode in (0, 1, 2):
            h = capi.DiscoHip(cfg, {**par, "C2P_Guess": mode, "C2P_Stop": 1}, state, library=host)
            if mode:
                h.set_c2p_model(trained[0])
            h.timestep(0.05)
            outs[mode], stats[mode] = h.download(a), h.c2p_stats()
            h.close()
        for mode in (1, 2):
            worst = max(field_errors(outs[mode], outs[0]).values())
            assert worst < 1e-12, (mode, worst)
        st = stats[2]
        assert st["learned"] + st["fallbacks"] == st["cells"]
        if poor:
            assert st["learned"] > 0.1 * st["cells"], st
            assert st["iterations"] <= 1.02 * min(stats[0]["iterations"],
                                                  stats[1]["iterations"]), stats
        else:
            assert st["learned"] < stats[1]["learned"], stats


def test_model_is_validated(host, trained):
    p = trained[0]
    cfg, par, a = _initial({"C2P_Guess": 1})
    g = capi.DiscoHip(cfg, par, a, library=host)
    with pytest.raises(capi.DhError, match="has no model"):
        g.timestep(0.01)
    with pytest.raises(capi.DhError, match="parameters"):
        g.set_c2p_model(p[:-1])
    bad = p.copy()
    bad[7] = np.nan
    with pytest.raises(capi.DhError, match="not finite"):
        g.set_c2p_model(bad)
    bad = p.copy()
    bad[c2p_model.N_IN] = 0.0  # in_sd[0]
    with pytest.raises(capi.DhError, match="in_sd"):
        g.set_c2p_model(bad)
    g.close()
    cfg, par, a = _initial()
    g = capi.DiscoHip(cfg, par, a, library=host)
    with pytest.raises(capi.DhError, match="not created with"):
        g.set_c2p_model(p)
    g.close()


def test_model_file_round_trip(tmp_path, trained):
    path = str(tmp_path / "m.f32")
    c2p_model.save(trained[0], path)
    assert os.path.getsize(path) == 4 * c2p_model.N_PARAM
    assert np.array_equal(c2p_model.load(path), trained[0].astype(np.float32))


def test_driver_reads_c2p_keys(tmp_path, trained):
    """The drop-in driver with in.par C2P_Guess 1, C2P_Stop 1 and C2P_Model
    runs from the model file and agrees with the plain driver to 1e-13; a
    missing model file is refused at set-up."""
    cfg = configs.get("gr_disk")
    exe = cpu_build.build_driver("gr_disk")
    model = str(tmp_path / "m.f32")
    c2p_model.save(trained[0], model)
    outs = []
    for extra in ({}, {"C2P_Guess": 1, "C2P_Stop": 1, "C2P_Model": model}):
        d = tmp_path / ("learned" if extra else "plain")
        state_io.write_par(str(d), cfg.par_dict(**PAR, **extra))
        subprocess.run([exe, "4"], cwd=d, check=True, capture_output=True, timeout=300)
        outs.append(state_io.read_state(str(d / "output.dstate")))
    assert max(field_errors(outs[1], outs[0]).values()) < 1e-13
    d = tmp_path / "missing"
    state_io.write_par(str(d), cfg.par_dict(**PAR, C2P_Guess=1, C2P_Model=str(tmp_path / "no")))
    r = subprocess.run([exe, "1"], cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode != 0 and "C2P_Model" in (r.stdout + r.stderr)


def test_shipped_model_guesses_the_disk(host):
    """The committed model (python -m disco_amd.c2p_model train) is used at
    every cell of the disk and, with the residual rule, needs at most one
    iteration more than the previous primitives."""
    if not os.path.exists(c2p_model.MODEL_PATH):
        pytest.skip("no shipped model")
    p = c2p_model.load()
    _, plain = _run(host, {"C2P_Stop": 1}, steps=2)
    out, stats = _run(host, {"C2P_Guess": 1, "C2P_Stop": 1}, p, steps=2)
    for s, r in zip(stats, plain):
        assert s["learned"] >= 0.999 * s["cells"], s
        assert s["iterations"] / s["cells"] <= r["iterations"] / r["cells"] + 1.05, (s, r)
