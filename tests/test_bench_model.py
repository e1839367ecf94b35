"""bench.py's measurement arithmetic on the CPU: the compulsory bytes per
zone-update, the roofline of the dominant kernel, the byte model of the
launches that small meshes merge, and the refusal of more than one GPU."""
import json
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
import bench  # noqa: E402


def test_b_alg_counts_40_bytes_per_state_component():
    # the state {cons (NQ), Phi (NF), piph} at 40 bytes per component per
    # zone-update: 560 for the MHD disk (8 + 5 + 1), 240 for Euler (5 + 1)
    assert bench.b_alg(8, 5) == 560
    assert bench.b_alg(5, 0) == 240


def test_dominant_roofline_uses_b_alg_over_the_mean_launch():
    kernels = {"a": {"ms_per_step": 2.0, "mean_launch_ms": 1.0, "model_bytes_per_launch": 5e9,
                     "model_GBps": 5000.0},
               "b": {"ms_per_step": 6.0, "mean_launch_ms": 3.0, "model_bytes_per_launch": 4e9,
                     "model_GBps": 1333.0}}
    r = bench.dominant_roofline(kernels, 560, 1_000_000, "no_such_workload")
    assert r["kernel"] == "b"
    assert r["algorithmic_bytes_per_launch"] == 280 * 1_000_000
    assert r["achieved"] == pytest.approx(280e6 / 3e-3 / 1e9)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0)
    assert r["share_of_step"] == pytest.approx(0.75)
    assert r["traffic"] is None


def test_merged_launches_model_is_the_sum_of_their_parts():
    args = (8, True, 1000, 800, 1900, 1800)
    kb = bench.kernel_bytes
    assert kb("flux_phi_r", *args) == kb("flux_phi", *args) + kb("flux_r", *args)
    assert kb("wcell_plm_phi", *args) == kb("wcell", *args) + kb("plm_phi", *args)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]
