"""The launch tuner (dh_autotune) on the GPU: a tuned run, whose kernels
take every traversal order in turn and then their fastest, equals an
untuned run bit for bit (every output is written by its owner, so the
order of the work changes nothing); the tuner keeps, for every kernel, the
variant with the least mean time of its sweep, and times every variant
when it has as many steps as the kernel has variants."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from disco_amd import capi, configs, state_io

pytestmark = pytest.mark.gpu


def _initial(cname, over):
    cfg = configs.get(cname)
    par = cfg.par_dict(**over)
    exe = configs.ref_binary(cfg)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    d = tempfile.mkdtemp(prefix="tuner_")
    state_io.write_par(d, par)
    subprocess.run([exe, d, "0"], cwd=d, check=True, capture_output=True)
    return cfg, par, state_io.read_state(os.path.join(d, "state_0000.bin"))


@pytest.mark.parametrize("cname,over", [
    ("mhd_kep", {"Num_R": 32, "Num_Z": 8}),
    ("euler_yee", {"Num_R": 64}),
])
def test_tuned_run_equals_untuned_and_keeps_the_fastest(cname, over):
    cfg, par, a = _initial(cname, over)
    print(f"case {cname} {over}: {a.n_real()} cells", flush=True)
    nvar = 2 + 2 * 4  # storage order (2) and up to four strip heights, twice
    tuned = capi.DiscoHip(cfg, par, a)
    tuned.autotune(nvar)
    plain = capi.DiscoHip(cfg, par, a)
    for n in range(nvar + 2):
        dt = plain.getmindt()
        assert tuned.getmindt() == dt, n
        tuned.timestep(dt)
        plain.timestep(dt)
    x, y = tuned.download(a), plain.download(a)
    for key in ("prim", "cons", "piph", "dphi", "phi_flux"):
        assert np.array_equal(getattr(x, key), getattr(y, key), equal_nan=True), key
    table = tuned.read_tuning()
    assert table, "no kernel was tuned"
    for name, row in table.items():
        sweep = row["ms_variants"]
        assert all(v >= 0.0 for v in sweep.values()), (name, sweep)
        assert row["ms_chosen"] == min(sweep.values()), (name, row)
        assert row["order"] in sweep, (name, row)
    tuned.close()
    plain.close()


def test_autotune_arguments():
    cfg, par, a = _initial("euler_yee", {"Num_R": 32})
    g = capi.DiscoHip(cfg, par, a)
    with pytest.raises(capi.DhError, match="negative"):
        g.autotune(-1)
    g.autotune(0)
    g.timestep(g.getmindt())
    g.close()
