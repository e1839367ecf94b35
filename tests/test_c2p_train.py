"""Training of the learned cons2prim guess with the engine's own kernels
(dh_c2p_grad, dh_c2p_train; disco_amd/csrc/dh_c2ptrain.h) on the host build,
against the float64 oracle (oracle/c2pnet.py), which is pinned here against
central finite differences."""
import os
import sys

import numpy as np
import pytest

from disco_amd import c2p_model, capi
from oracle import c2pnet

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpu_engine"))
import build as cpu_build  # noqa: E402


def _problem(n=300, seed=3):
    """A random model and rows whose targets are a smooth function of the
    features (so that training can reduce the error)."""
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 8)).astype(np.float32) * 2.0 + 1.0
    Y = np.stack([np.tanh(X[:, 0] - 0.5 * X[:, 1]), 0.3 * X[:, 2] * X[:, 3] + X[:, 4]],
                 axis=1).astype(np.float32)
    p = np.zeros(c2pnet.N_PARAM, np.float32)
    for k in ("W1", "b1", "W2", "b2", "W3", "b3"):
        o, m = c2pnet.OFF[k]
        fan_in = 8 if k in ("W1", "b1") else 64
        p[o:o + m] = rng.uniform(-1, 1, m) / np.sqrt(fan_in)
    p[c2pnet.OFF["in_mu"][0]:][:8] = X.mean(0)
    p[c2pnet.OFF["in_sd"][0]:][:8] = X.std(0)
    p[c2pnet.OFF["out_mu"][0]:][:2] = Y.mean(0)
    p[c2pnet.OFF["out_sd"][0]:][:2] = Y.std(0)
    return p, X, Y


def test_oracle_gradient_matches_finite_differences():
    p, X, Y = _problem(n=20)
    g, _ = c2pnet.grad(p, X, Y)
    rng = np.random.default_rng(0)
    trainable = np.arange(c2pnet.OFF["W1"][0], c2pnet.OFF["out_mu"][0])
    p64 = p.astype(np.float64)
    for q in rng.choice(trainable, 60, replace=False):
        h = 1e-6
        a, b = p64.copy(), p64.copy()
        a[q] += h
        b[q] -= h
        fd = (c2pnet.loss(a, X, Y) - c2pnet.loss(b, X, Y)) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(fd)), (q, fd, g[q])
    # normalisation constants are not trained
    assert not g[:c2pnet.OFF["W1"][0]].any() and not g[c2pnet.OFF["out_mu"][0]:].any()


@pytest.fixture(scope="module")
def host():
    return capi.load_library(cpu_build.build())


def test_engine_gradient_matches_oracle(host):
    p, X, Y = _problem()
    g, loss = c2p_model.grad(p, X, Y, library=host)
    go, lo = c2pnet.grad(p, X, Y)
    assert abs(loss - lo) <= 1e-5 * lo
    scale = np.abs(go).max()
    assert np.abs(g - go).max() <= 2e-5 * scale, np.abs(g - go).max() / scale


def test_engine_training_reduces_the_error(host):
    p, X, Y = _problem(n=2000)
    _, l1 = c2p_model.fit(X, Y, epochs=1, batch=128, lr=3e-3, seed=5, library=host)
    p10, l10 = c2p_model.fit(X, Y, epochs=10, batch=128, lr=3e-3, seed=5, library=host)
    assert l10 < 0.5 * l1, (l1, l10)
    # the normalisation is the data's, and the returned error is the model's
    assert np.allclose(p10[c2pnet.OFF["in_mu"][0]:][:8], X.mean(0), rtol=1e-5, atol=1e-6)
    assert abs(c2pnet.loss(p10, X, Y) - l10) < 0.5 * l10
    assert np.all(np.isfinite(p10))
