"""The learned initial guess of grmhd2's cons2prim: training data, training
and the model file.

    python -m disco_amd.c2p_model train [--out FILE] [--steps N] [--size NRxNZ]

noble2d (grmhd2.c:1096-1314) solves for (v^2, eta) given four invariants of
the conserved state; the model (disco_amd/csrc/dh_c2pnet.h: 8 -> 64 -> 64 ->
2, ReLU, bfloat16 weights on MFMA) maps features of those invariants to the
solution.  Training data come from GR disk runs through the host layer
(`harvest`, dh_c2p_harvest after every step); `fit` trains the MLP with the
engine's own float32 forward, backward and Adam kernels (dh_c2p_train) and
returns the parameter block; `save`/`load` keep it as raw little-endian
float32, the file the host layer's in.par key C2P_Model names.
"""
from __future__ import annotations

import argparse
import ctypes as C
import os
import sys
import tempfile
import time
from typing import Iterable, List, Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(HERE, "models", "c2p_grmhd.f32")

N_IN, N_H, N_OUT = 8, 64, 2
N_PARAM = 2 * N_IN + N_IN * N_H + N_H + N_H * N_H + N_H + N_H * N_OUT + N_OUT + 2 * N_OUT

# GR disk runs of the training set: (config, in.par overrides) on r in
# [8, 20], outside the innermost stable orbit (6 M); every solver, both
# metrics, weak and strong field, static and moving meshes.  Each stays
# finite over 30 steps (with HLLD on the moving mesh at CFL 0.2).
TRAINING_BASE = {"R_Min": 8.0, "R_Max": 20.0}
TRAINING_RUNS = [
    ("gr_disk", {"Riemann_Solver": 2, "CFL": 0.2}),
    ("gr_disk", {"Riemann_Solver": 2, "Mesh_Motion": 0}),
    ("gr_disk", {"Riemann_Solver": 2, "Init_Par1": 10.0, "CFL": 0.2}),
    ("gr_disk", {"Riemann_Solver": 0}),
    ("gr_disk", {"Riemann_Solver": 1, "Mesh_Motion": 0, "Init_Par0": 1}),
    ("gr_disk3d", {"Riemann_Solver": 0}),
]


def physical(f: np.ndarray, t: np.ndarray) -> np.ndarray:
    """Rows whose state is one the model should learn: finite, v^2 below
    0.999 and non-negative field invariants (features 2 and 3: Q / e and
    psi / e^3)."""
    return (np.isfinite(f).all(axis=1) & np.isfinite(t).all(axis=1) & (t[:, 0] >= 0.0)
            & (t[:, 0] < 0.999) & (f[:, 2] >= 0.0) & (f[:, 3] >= 0.0))


def load(path: str = MODEL_PATH) -> np.ndarray:
    p = np.fromfile(path, dtype="<f4")
    if p.size != N_PARAM:
        raise ValueError(f"{path}: {p.size} parameters, expected {N_PARAM}")
    return p


def save(params: np.ndarray, path: str = MODEL_PATH) -> None:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.asarray(params, dtype="<f4").ravel().tofile(path)


def harvest_ctx(lib, ctx, ntot: int) -> Tuple[np.ndarray, np.ndarray]:
    """Features and targets of the real cells of a device context."""
    f = np.zeros((ntot, N_IN), np.float32)
    t = np.zeros((ntot, N_OUT), np.float32)
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    if lib.dh_c2p_harvest(ctx, fp(f), fp(t)):
        raise RuntimeError(f"dh_c2p_harvest: {lib.dh_last_error().decode()}")
    keep = physical(f, t)
    return f[keep], t[keep]


def harvest(runs: Iterable[Tuple[str, dict]], steps: int, size: Tuple[int, int],
            per_step: int = 5000, seed: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    """Runs each GR disk through the host layer on the GPU and keeps up to
    per_step random real cells' (features, targets) after every step."""
    from . import capi, configs, state_io
    from .host_lib import HostRun
    rng = np.random.default_rng(seed)
    lib = capi.lib()
    xs: List[np.ndarray] = []
    ys: List[np.ndarray] = []
    for name, over in runs:
        cfg = configs.get(name)
        par = cfg.par_dict(Num_R=size[0], Num_Z=size[1], **{**TRAINING_BASE, **over})
        d = tempfile.mkdtemp(prefix="c2ph_")
        run = HostRun(name, state_io.write_par(d, par))
        ctx, ntot = C.c_void_p(run.ctx), run.ntot
        for n in range(steps):
            run.timestep(run.getmindt())
            f, t = harvest_ctx(lib, ctx, ntot)
            if len(f) > per_step:
                pick = rng.choice(len(f), per_step, replace=False)
                f, t = f[pick], t[pick]
            xs.append(f)
            ys.append(t)
        run.close()
        print(f"harvest {name} {over}: {steps} steps, {sum(len(x) for x in xs)} samples so far",
              flush=True)
    return np.concatenate(xs), np.concatenate(ys)


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def fit(X: np.ndarray, Y: np.ndarray, epochs: int = 40, batch: int = 2048, lr: float = 3e-3,
        seed: int = 0, library=None) -> Tuple[np.ndarray, float]:
    """Trains the model on (features, targets) with the engine's own
    training kernels (dh_c2p_train: on the GPU, or on the CPU with the host
    build as `library`); returns the parameter block and the last epoch's
    mean squared error (normalised targets)."""
    from . import capi
    lib = library if library is not None else capi.lib()
    X = np.ascontiguousarray(X, np.float32)
    Y = np.ascontiguousarray(Y, np.float32)
    p = np.zeros(N_PARAM, np.float32)
    loss = C.c_float()
    if lib.dh_c2p_train(_fp(X), _fp(Y), len(X), epochs, batch, lr, seed, _fp(p), C.byref(loss)):
        raise RuntimeError(f"dh_c2p_train: {lib.dh_last_error().decode()}")
    return p, loss.value


def grad(params: np.ndarray, X: np.ndarray, Y: np.ndarray, library=None):
    """(gradient of the mean squared error over all rows with respect to the
    parameter block, the error) from the engine's kernels (dh_c2p_grad)."""
    from . import capi
    lib = library if library is not None else capi.lib()
    params = np.ascontiguousarray(params, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    Y = np.ascontiguousarray(Y, np.float32)
    g = np.zeros(N_PARAM, np.float32)
    loss = C.c_float()
    if lib.dh_c2p_grad(_fp(params), _fp(X), _fp(Y), len(X), _fp(g), C.byref(loss)):
        raise RuntimeError(f"dh_c2p_grad: {lib.dh_last_error().decode()}")
    return g, loss.value


def predict(params: np.ndarray, X: np.ndarray) -> np.ndarray:
    """The model's targets in float64 (no bfloat16 rounding), for checks."""
    p = np.asarray(params, np.float64)
    o = 0
    def take(n):  # noqa: E306
        nonlocal o
        a = p[o:o + n]
        o += n
        return a
    in_mu, in_sd = take(N_IN), take(N_IN)
    W1, b1 = take(N_IN * N_H).reshape(N_IN, N_H), take(N_H)
    W2, b2 = take(N_H * N_H).reshape(N_H, N_H), take(N_H)
    W3, b3 = take(N_H * N_OUT).reshape(N_H, N_OUT), take(N_OUT)
    out_mu, out_sd = take(N_OUT), take(N_OUT)
    x = (np.asarray(X, np.float64) - in_mu) / in_sd
    h = np.maximum(x @ W1 + b1, 0.0)
    h = np.maximum(h @ W2 + b2, 0.0)
    return (h @ W3 + b3) * out_sd + out_mu


def report(params: np.ndarray, X: np.ndarray, Y: np.ndarray) -> dict:
    """Errors of the model's guess on (X, Y): |dv^2| and |d eta / eta|."""
    P = predict(params, X)
    dv2 = np.abs(P[:, 0] - Y[:, 0])
    deta = np.abs(np.expm1(P[:, 1] - Y[:, 1]))
    q = lambda a: {"median": float(np.median(a)), "p99": float(np.quantile(a, 0.99)),  # noqa: E731
                   "max": float(a.max())}
    return {"samples": int(len(X)), "dv2": q(dv2), "deta_rel": q(deta)}


def main(argv: Optional[List[str]] = None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    tr = sub.add_parser("train")
    tr.add_argument("--out", default=MODEL_PATH)
    tr.add_argument("--steps", type=int, default=30)
    tr.add_argument("--size", default="64x16")
    tr.add_argument("--epochs", type=int, default=40)
    a = ap.parse_args(argv)
    nr, nz = (int(v) for v in a.size.split("x"))
    t0 = time.time()
    X, Y = harvest(TRAINING_RUNS, a.steps, (nr, nz))
    print(f"harvested {len(X)} samples in {time.time() - t0:.1f} s", flush=True)
    rng = np.random.default_rng(1)
    test = rng.random(len(X)) < 0.1
    t0 = time.time()
    p, loss = fit(X[~test], Y[~test], epochs=a.epochs)
    print(f"trained on the GPU in {time.time() - t0:.1f} s: last-epoch mse {loss:.3e}", flush=True)
    print("held-out errors:", report(p, X[test], Y[test]), flush=True)
    save(p, a.out)
    print(f"wrote {a.out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
