"""Every grmhd2 cons2prim residual against float64 prim2cons: the conserved
state that the recovered primitives give back, through the reference's own
prim2cons (grmhd2.c, built into oracle/_ref and run on a state file by its
driver's p2c mode), must match the conserved state cons2prim started from,
cell by cell.

Excluded are the cells where the primitives are not meant to reproduce the
conserved state: a floor set them (density at Density_Floor, or pressure at
Pressure_Floor times density, grmhd2.c's floors), or the Newton iteration
did not meet its stopping rule within its 30 iterations (the record
dh_c2p_info keeps; the reference's own solver fails there as well, the host
build being bit-exact with it).  These are a few per cent of the torus's
cells at most, at its atmosphere's edge; the test bounds their share.

The relative error of a conserved component is |U_ref - U| / (|U| + |D|),
D the conserved density: 1e-10 is the bar, measured values are 1e-15.

The CPU leg checks the check itself on the reference build's own output;
the GPU legs run the device's steps, with the reference's stopping rule and
with the residual rule, and with the learned guess."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from disco_amd import capi, configs, state_io

RES_TOL = 1e-10
MAX_EXCLUDED = 0.05
C2P_FAILED = 63


def _run_ref(cname, over, nsteps):
    cfg = configs.get(cname)
    par = cfg.par_dict(**over)
    exe = configs.ref_binary(cfg)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    d = tempfile.mkdtemp(prefix="c2pres_")
    state_io.write_par(d, {k: v for k, v in par.items() if k not in state_io.EXTRA_KEYS})
    subprocess.run([exe, d, str(nsteps)], cwd=d, check=True, capture_output=True)
    return cfg, par, d


def ref_prim2cons(cfg, d, s):
    """prim2cons of the reference build for every real cell of state s (run
    in directory d, which holds the run's in.par)."""
    state_io.write_state(os.path.join(d, "p2c_in.bin"), s)
    subprocess.run([configs.ref_binary(cfg), "p2c", "p2c_in.bin", "p2c_out.bin"], cwd=d,
                   check=True, capture_output=True)
    return np.fromfile(os.path.join(d, "p2c_out.bin")).reshape(s.cons.shape)


def residuals(cfg, par, d, s, failed=None):
    """(relative error per real cell, mask of the cells checked)."""
    real = s.real_mask()
    cref = ref_prim2cons(cfg, d, s)
    c, cr = s.cons[:, real], cref[:, real]
    err = (np.abs(cr - c) / (np.abs(c) + np.abs(c[0]))).max(axis=0)
    rho, prs = s.prim[0, real], s.prim[1, real]
    floored = (rho == par["Density_Floor"]) | (prs == par["Pressure_Floor"] * rho)
    keep = ~floored
    if failed is not None:
        keep &= ~failed[real]
    return err, keep


def test_reference_output_passes_its_own_prim2cons():
    """The reference build's own state after 3 steps: its cons2prim output
    reproduces its conserved state through its prim2cons (floors and
    unconverged solves aside), so the p2c path and the error measure are
    sound."""
    cfg, par, d = _run_ref("kerr_torus", {"Num_R": 24, "Num_Z": 12}, 3)
    s = state_io.read_state(os.path.join(d, "state_final.bin"))
    err, keep = residuals(cfg, par, d, s)
    # the reference keeps no record of unconverged solves: allow the few
    # cells they leave, and hold the rest to the bar
    bad = err[keep] > RES_TOL
    assert bad.mean() < MAX_EXCLUDED, bad.mean()
    assert np.quantile(err[keep], 0.9) < 1e-14


def test_planted_error_in_the_primitives_is_seen():
    """A primitive moved by 1e-8 (relative) in one cell shows up as a
    residual above the bar: the check is not blind."""
    cfg, par, d = _run_ref("gr_disk", {"Num_R": 16, "Num_Z": 4, "Riemann_Solver": 2}, 2)
    s = state_io.read_state(os.path.join(d, "state_final.bin"))
    err, keep = residuals(cfg, par, d, s)
    assert err[keep].max() <= RES_TOL
    real = np.nonzero(s.real_mask())[0]
    s.prim[2, real[len(real) // 2]] *= 1.0 + 1e-8
    err2, keep2 = residuals(cfg, par, d, s)
    assert err2[keep2].max() > 10 * RES_TOL


GPU_CASES = [
    ("gr_disk", {"Num_R": 48, "Num_Z": 12, "Riemann_Solver": 2}, {}),
    ("gr_disk", {"Num_R": 48, "Num_Z": 12, "Riemann_Solver": 2}, {"C2P_Stop": 1}),
    ("kerr_torus", {"Num_R": 48, "Num_Z": 24}, {}),
    ("kerr_torus", {"Num_R": 48, "Num_Z": 24}, {"C2P_Stop": 1}),
]


@pytest.mark.gpu
@pytest.mark.parametrize("cname,over,extra", GPU_CASES)
def test_gpu_cons2prim_residuals(cname, over, extra):
    cfg, par, d = _run_ref(cname, over, 0)
    a = state_io.read_state(os.path.join(d, "state_0000.bin"))
    par = {**par, **extra}
    print(f"case {cname} {over} {extra}: {a.n_real()} cells, 4 steps", flush=True)
    g = capi.DiscoHip(cfg, par, a)
    for _ in range(4):
        g.timestep(g.getmindt())
    s = g.download(a)
    info = g.c2p_info()
    g.close()
    err, keep = residuals(cfg, par, d, s, failed=(info & C2P_FAILED) == C2P_FAILED)
    assert 1.0 - keep.mean() < MAX_EXCLUDED, 1.0 - keep.mean()
    assert err[keep].max() <= RES_TOL, err[keep].max()


@pytest.mark.gpu
def test_gpu_learned_guess_cons2prim_residuals():
    """With the learned initial guess (shipped model) and the residual rule."""
    from disco_amd import c2p_model
    cfg, par, d = _run_ref("gr_disk", {"Num_R": 48, "Num_Z": 12, "Riemann_Solver": 2}, 0)
    a = state_io.read_state(os.path.join(d, "state_0000.bin"))
    par = {**par, "C2P_Stop": 1, "C2P_Guess": 1}
    g = capi.DiscoHip(cfg, par, a)
    g.set_c2p_model(c2p_model.load(c2p_model.MODEL_PATH))
    for _ in range(4):
        g.timestep(g.getmindt())
    s = g.download(a)
    info = g.c2p_info()
    g.close()
    assert ((info & 64) != 0).sum() > 0, "the learned guess was never used"
    err, keep = residuals(cfg, par, d, s, failed=(info & C2P_FAILED) == C2P_FAILED)
    assert 1.0 - keep.mean() < MAX_EXCLUDED
    assert err[keep].max() <= RES_TOL, err[keep].max()
