"""CPU oracle of the learned cons2prim guess's training: the MLP's loss and
its gradient in float64 (numpy), for checking the engine's float32 forward,
backward and gradient kernels (disco_amd/csrc/dh_c2ptrain.h).

Test infrastructure only: tests/ import it as the checker; the product never
does.  It restates the mathematics of the model (no reference counterpart:
the learned guess is new), and is itself pinned against central finite
differences of `loss` (tests/test_c2p_train.py).

Parameter block layout (float32 in the engine, float64 here):
  in_mu[8] in_sd[8] W1[8][64] b1[64] W2[64][64] b2[64] W3[64][2] b3[2]
  out_mu[2] out_sd[2];  W[k][j] multiplies input k into output j.
Loss: mean over rows and the two outputs of (y - t)^2 with normalised inputs
x = (X - in_mu) / in_sd and targets t = (Y - out_mu) / out_sd; ReLU hidden
layers, whose derivative at 0 is 0.
"""
import numpy as np

N_IN, N_H, N_OUT = 8, 64, 2
OFF = {}
_o = 0
for _name, _n in (("in_mu", N_IN), ("in_sd", N_IN), ("W1", N_IN * N_H), ("b1", N_H),
                  ("W2", N_H * N_H), ("b2", N_H), ("W3", N_H * N_OUT), ("b3", N_OUT),
                  ("out_mu", N_OUT), ("out_sd", N_OUT)):
    OFF[_name] = (_o, _n)
    _o += _n
N_PARAM = _o


def unpack(p):
    p = np.asarray(p, np.float64)
    get = lambda k: p[OFF[k][0]:OFF[k][0] + OFF[k][1]]  # noqa: E731
    return {"in_mu": get("in_mu"), "in_sd": get("in_sd"),
            "W1": get("W1").reshape(N_IN, N_H), "b1": get("b1"),
            "W2": get("W2").reshape(N_H, N_H), "b2": get("b2"),
            "W3": get("W3").reshape(N_H, N_OUT), "b3": get("b3"),
            "out_mu": get("out_mu"), "out_sd": get("out_sd")}


def loss(p, X, Y):
    q = unpack(p)
    x = (np.asarray(X, np.float64) - q["in_mu"]) / q["in_sd"]
    t = (np.asarray(Y, np.float64) - q["out_mu"]) / q["out_sd"]
    h1 = np.maximum(x @ q["W1"] + q["b1"], 0.0)
    h2 = np.maximum(h1 @ q["W2"] + q["b2"], 0.0)
    y = h2 @ q["W3"] + q["b3"]
    return float(np.mean((y - t) ** 2))


def grad(p, X, Y):
    """(d loss / d parameter block, loss); zero at the normalisation slots."""
    q = unpack(p)
    x = (np.asarray(X, np.float64) - q["in_mu"]) / q["in_sd"]
    t = (np.asarray(Y, np.float64) - q["out_mu"]) / q["out_sd"]
    n = len(x)
    z1 = x @ q["W1"] + q["b1"]
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ q["W2"] + q["b2"]
    h2 = np.maximum(z2, 0.0)
    y = h2 @ q["W3"] + q["b3"]
    d3 = 2.0 * (y - t) / (N_OUT * n)
    d2 = (d3 @ q["W3"].T) * (z2 > 0)
    d1 = (d2 @ q["W2"].T) * (z1 > 0)
    g = np.zeros(N_PARAM)
    put = lambda k, v: g.__setitem__(slice(OFF[k][0], OFF[k][0] + OFF[k][1]), np.ravel(v))  # noqa: E731
    put("W1", x.T @ d1)
    put("b1", d1.sum(0))
    put("W2", h1.T @ d2)
    put("b2", d2.sum(0))
    put("W3", h2.T @ d3)
    put("b3", d3.sum(0))
    return g, float(np.mean((y - t) ** 2))
