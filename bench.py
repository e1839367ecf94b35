#!/usr/bin/env python3
"""Throughput of the Disco timestep on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one full RK2 timestep (getmindt + timestep, both through the
reference-signature host entry points of libdisco_host_<config>.so) over the
whole mesh, with the state resident in HBM.  The default workload is the 3-D
MHD Keplerian disk (Nr=256, Nz=64, constrained transport).  The timed region
is exactly K steps between device synchronisations.  Prints one JSON line:

  value          zone-updates/s = real cells x K / seconds
  roofline       the dominant kernel (most device time per step) against the
                 8 TB/s HBM3E peak: achieved = B_alg / 2 x real cells (the
                 state bytes of one substep, which one launch of a per-substep
                 kernel advances) / its mean launch time (HIP events on the
                 library's stream); traffic is the measured HBM bytes per
                 launch of that kernel from the rocprofv3 counter summary under
                 profiles/, when there is one.  kernel_model_GBps beside it is
                 the kernel's own byte model (every array it touches read or
                 written once) over the same time
  step_roofline  B_alg x zone-updates/s against 8 TB/s, with
                 B_alg = 40 (NQ + NF + 1) bytes per zone-update (the state
                 {cons, Phi, piph} read and written over the RK2 step); also
                 against the bandwidth a plain device copy measures on this
                 GPU (measured_copy), as roofline is
  grmhd_c2p      the GRMHD workload stepped with the plain (previous
                 primitives) guess, the learned one and the better of the
                 two per cell, all with the residual stopping rule (both
                 residuals below 1e-10): the ratios of their zone-updates/s
  cpu_baseline   the reference build (oracle/_ref) on this host's cores: one
                 serial process per core, each stepping the same bounded
                 sample of the workload; value = their summed zone-updates/s
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

from disco_amd import configs, state_io  # noqa: E402

PEAK_HBM_GBS = 8000.0

WORKLOADS = {
    "mhd_kep": dict(
        config="mhd_kep", over={}, steps=20, warmup=2,
        desc="3-D Newtonian MHD Keplerian disk (MRI seed): r in [0.5,2], z in [-0.25,0.25], "
             "Nr=256, Nz=64, Np=int(2 pi r/dr) per annulus, rho=1, cs=v_K/20, vertical field "
             "at beta=1000, velocity noise 5e-5 v_phi; HLLD, PLM 1.5, constrained transport, "
             "Keplerian mesh motion, RK2, CFL 0.5",
        cpu_over={"Num_R": 64, "Num_Z": 16}, cpu_steps=4),
    "euler_yee": dict(
        config="euler_yee", over={}, steps=100, warmup=5,
        desc="2-D Euler isentropic vortex, Nr=128 polar mesh (51,809 cells), HLLC, PLM 1.5, "
             "RK2, CFL 0.5, rigid mesh rotation",
        cpu_over={}, cpu_steps=150),
    "mhd_c5": dict(
        config="mhd_kep",
        over={**configs.C5_PAR, "Num_R": 512, "Num_Z": 160},
        steps=5, warmup=1,
        desc="3-D Newtonian MHD Keplerian disk at the one-GPU memory-filling size: r in [0.5,3], "
             "z in [-0.3,0.3], Nr=512, Nz=160, Np=int(2 pi r/dr) per annulus, rho=1, "
             "cs=v_K/20, zero-net-flux toroidal field B_phi ~ sin(2 pi (r-0.5)/2.5) at beta=100, "
             "velocity noise 5e-5 v_phi; HLLD, PLM 1.5, constrained transport, Keplerian mesh "
             "motion, RK2, CFL 0.5",
        cpu_over={**configs.C5_PAR, "Num_R": 64, "Num_Z": 20},
        cpu_steps=3),
    "disk_ensemble": dict(
        config="disk_planet", over={}, steps=200, warmup=2, members=64, seed=314159,
        desc="64 independent 2-D isothermal disk-planet Euler runs, each its own mesh, dt and "
             "planet, stepped together: r in [0.4,2.5], Nr drawn from {64,128,256,512} by seed, "
             "Np=int(2 pi r/dr), Sigma=r^-0.5, h=0.05, alpha viscosity 1e-3, a planet of mass "
             "ratio q drawn log-uniform in [1e-5,1e-3] on a circular orbit at r=1 softened by "
             "0.6h; HLLC, PLM 1.5, RK2, CFL 0.5, Keplerian mesh motion",
        cpu_over={"Num_R": 128}, cpu_steps=60),
    "gr_disk": dict(
        config="gr_disk",
        over={"Num_R": 128, "Num_Z": 32, "R_Min": 8.0, "R_Max": 20.0, "Riemann_Solver": 2,
              "CFL": 0.2},
        steps=10, warmup=1,
        desc="3-D GRMHD disk on the Schwarzschild equator (Kerr-Schild, M=1): r in [8,20], "
             "z in [-0.5,0.5], Nr=128, Nz=32, circular-geodesic gas rho=(r/6)^-1.5, "
             "P=0.01 rho, gamma 4/3, vertical field at beta=100; relativistic HLLD, PLM 1.5, "
             "constrained transport, Keplerian mesh motion, RK2, CFL 0.2",
        cpu_over={"Num_R": 32, "Num_Z": 8, "R_Min": 8.0, "R_Max": 20.0, "Riemann_Solver": 2,
                  "CFL": 0.2},
        cpu_steps=10),
    "kerr_torus": dict(
        config="kerr_torus", over={}, steps=10, warmup=2,
        desc="3-D GRMHD Fishbone-Moncrief torus around a Kerr black hole (a=0.9, M=1) in "
             "Kerr-Schild cylindrical coordinates: r in [2,40], z in [-20,20], Nr=192, Nz=96, "
             "aspect 2.5 (Np up to 512), torus inner edge r=6, pressure maximum r=12, "
             "gamma 4/3, poloidal loops on density contours at beta=100 plus a uniform "
             "B^z=1e-4, 1% pressure noise, atmosphere rho 1e-2, P 1e-5; relativistic HLLD, "
             "PLM 1.5, constrained transport, mesh moving with the gas, RK2, CFL 0.05",
        cpu_over={"Num_R": 48, "Num_Z": 24}, cpu_steps=4),
}
DEFAULT_WORKLOAD = "mhd_kep"


def b_alg(nq: int, nf: int) -> int:
    return 40 * (nq + nf + 1)


def kernel_bytes(name: str, nq: int, mhd: bool, nt: int, nr: int, nfr: int, nfz: int,
                 zlayer: int = 0) -> float:
    """Algorithmic bytes of one launch of a kernel: every array element the
    kernel needs read once and every output element written once (8-byte
    doubles, 4-byte indices, 1-byte face orientations).  nt counts all cells
    (ghosts included), nr the real ones, nfr/nfz the r and z faces; zlayer is
    the cell count of one z ghost layer (0 when no z side is copied).

    Per face the topology holds L, R (int), LR (byte) and up, dA, phic, dphi
    (double); every cell has two merge steps per direction (int) and its
    annulus index cjk (int)."""
    merged = {"wcell_plm_phi": ("wcell", "plm_phi"), "flux_phi_r": ("flux_phi", "flux_r")}
    if name in merged:                  # two launches run as one on small meshes
        return sum(kernel_bytes(k, nq, mhd, nt, nr, nfr, nfz, zlayer) for k in merged[name])
    d, i = 8, 4
    nql = nq - 3 if mhd else nq         # components the fluxes update
    nf = {"r": nfr, "z": nfz}.get(name[-1:], 0)
    nfa = nfr + nfz
    steps = 2 * i * nt                  # the two merge steps of one direction
    if name == "wcell":                 # analytic mesh motion: r_jph in, wiph out
        return d * nt + i * nt
    if name == "ring_start":
        return d * nt
    if name == "plm_phi":
        return d * (2 * nq * nt + nt) + i * nt
    if name in ("ranges_r", "ranges_z"):
        return d * nt + i * nt + steps
    if name in ("faces_r", "faces_z"):
        return steps + i * nt + d * 2 * nt + nf * (2 * i + 1 + 4 * d)
    if name in ("grad_r", "grad_z"):
        return steps + i * nt + nf * (2 * i + 2 * d) + d * nt * (2 + 3 * nq)
    if name == "flux_phi":
        return d * (2 * nq + nql + 2) * nt + i * nt + (d * 4 * nt if mhd else 0)
    if name in ("flux_r", "flux_z"):
        return (nf * (2 * i + 2 * d) + d * nt * (2 + 3 * nq) + d * nql * nf
                + (3 * d * nf if mhd else 0))
    if name == "edges_collect":
        return d * 21 * nt + 2 * steps + i * nt + 3 * d * nfa
    if name == "edges_final":
        # Epre (16), wiph, piph and dphi in, Eo and Eown (4 + 4) out per cell;
        # per owned face its cells and orientation and its E_phi in
        return d * 27 * nt + 2 * steps + i * nt + nfa * (d + 2 * i + 1)
    if name == "update_B":
        # Phi in and out, rkPhi on the second substep only (mean 2.5 per cell)
        return d * (12.5 + 4) * nt + 2 * steps + i * nt + nfa * (2 * i + 1)
    if name in ("line_steps_r", "line_steps_z"):
        return nf * (d + 2 * i)
    if name == "edge_segments":
        # face list, merge steps and line offsets in, the four cells' phi
        # edges and E_phi in, one E dl dt per line step (about two per face) out
        return nfa * (2 * i + 2 * i) + d * 6 * nt + d * 2 * nfa
    if name == "edge_adjust":
        # face list and merge steps in, the stored segments in, Phi in and out
        return nfa * (2 * i + 1 + 2 * i + 2 * d) + d * 4 * nfa
    if name == "update":
        # cons in/out, rkcons on the second substep only, prim in for the
        # sources (+ prim out for hydro), phi/r/z face deltas, geometry in
        # and moved geometry out; with constrained transport the field
        # components of cons and of the face deltas are not touched
        return (d * nr * (2.5 * nql + nq + (0 if mhd else nq)) + d * nql * (nt + nfa)
                + 2 * steps + i * nt + d * 5 * nt)
    if name == "c2p_guess":            # the invariants' inputs in, (v^2, eta) out
        return d * 8 * nr + d * 2 * nt + i * nt + 4 * 2 * nt
    if name == "c2p":                  # cons and prim in, guess in, prim and record out
        return d * nq * nr * 3 + d * 2 * nt + i * nt + 4 * 2 * nt + nt
    if name == "bfaces_c2p":
        return (2 * steps + i * nt + d * 10 * nt + nfa * (2 * i + 1 + d)
                + d * nq * nr * 2)
    if name in ("flip_flags_r", "flip_flags_z"):
        return nf * (2 * i + 2) + 2 * d * nt
    if name == "flips":
        return nfa
    if name == "mindt":
        return d * (nq * nr + 3 * nt) + i * nt
    if name == "diagnostics":          # prim (rho..B), piph, dphi in; Qrz out
        return d * (8 + 2) * nt + i * nt
    if name == "bc_scan":               # one z side: positions of both layers' rings
        return 2 * zlayer * (4 * d + i)
    if name == "bc_fill":               # one z ghost layer from the layer inside it
        return zlayer * (d * (2 * nq + 4) + i)
    if name == "copy":                  # cons -> rkcons (and Phi -> rkPhi): mean launch
        return 2 * d * nt * (nq + 5) / 2 if mhd else 2 * d * nt * nq
    return 0.0


def host_cpu():
    """(model name, logical CPUs this process may run on) of the host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    return model, ncpu


# The GPU box gives one GPU's share of its host: 16 cores.
CPU_BASELINE_MAX_CORES = 16


def cpu_baseline(cfg, over, steps, cores=None):
    """The reference build (oracle/_ref, serial: USE_MPI=0) on this host:
    `cores` processes, one per core, each stepping the same sample of the
    workload; value = the sum of their zone-updates/s (cells x steps over
    each process's own stepping time), i.e. the host's throughput with the
    work split into independent pieces, an upper bound for the reference's
    MPI decomposition of the same cells."""
    exe = configs.ref_binary(cfg)
    if not os.path.exists(exe):
        return None
    model, ncpu = host_cpu()
    cores = cores or min(ncpu, CPU_BASELINE_MAX_CORES)
    par = cfg.par_dict(**over)
    with tempfile.TemporaryDirectory(prefix="bench_ref_") as d:
        procs = []
        for i in range(cores):
            wd = os.path.join(d, str(i))
            state_io.write_par(wd, par)
            procs.append(subprocess.Popen([exe, wd, str(steps)], cwd=wd, stdout=subprocess.PIPE,
                                          stderr=subprocess.DEVNULL, text=True))
        results = []
        for p in procs:
            out, _ = p.communicate(timeout=900)
            if p.returncode != 0:
                return None
            line = [ln for ln in out.splitlines() if ln.startswith("{")][-1]
            results.append(json.loads(line))
    value = sum(r["zone_updates_per_s"] for r in results)
    secs = max(r["seconds"] for r in results)
    size = ", ".join(f"{k}={v}" for k, v in over.items()) or "the full mesh"
    return {"value": value, "unit": "zone-updates/s", "cores": cores, "kind": "reference",
            "cpu_model": model, "host_logical_cpus": ncpu,
            "per_core": value / cores,
            "sample": f"{cores} concurrent serial runs (one per core) of {steps} RK2 steps of the "
                      f"same workload at {size} ({results[0]['cells']} cells each) with the "
                      f"reference timestep built from its sources (gcc -O3, USE_MPI=0); "
                      f"{secs:.1f} s of stepping; value is their summed zone-updates/s"}


def _copy_bandwidth_here(gib: int = 4):
    """A gib-GiB float64 buffer copied device to device 10 times after one
    untimed copy, timed with events on torch's stream: bytes read and
    written over the time."""
    import torch
    if not torch.cuda.is_available():
        return None
    n = gib << 27
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 10
    e0.record()
    for _ in range(reps):
        y.copy_(x)
    e1.record()
    e1.synchronize()
    secs = e0.elapsed_time(e1) * 1e-3 / reps
    return {"GBps": 2 * n * 8 / secs / 1e9, "bytes_per_copy": 2 * n * 8,
            "how": f"torch device-to-device copy of {gib} GiB, mean of {reps}, own process"}


def copy_bandwidth():
    """The HBM bandwidth a plain device-to-device copy reaches on this GPU,
    measured in a child process (torch's runtime set up beside this
    process's library would not see the device); None if it cannot be."""
    code = ("import json, sys; sys.path.insert(0, %r); import bench; "
            "print(json.dumps(bench._copy_bandwidth_here()))" % REPO)
    try:
        r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                           timeout=300)
        return json.loads(r.stdout.strip().splitlines()[-1]) if r.returncode == 0 else None
    except (subprocess.TimeoutExpired, ValueError, IndexError):
        return None


def dominant_roofline(kernels, ba, ncell, workload):
    """Roofline of the kernel with the most device time per step.  Its
    algorithmic bytes per launch are the compulsory per-zone-update figure
    B_alg taken over one substep (B_alg / 2: the RK2 step reads and writes
    the state {cons, Phi, piph} twice) times the real cells a launch covers;
    the kernel's own byte model (every array it touches once) is reported
    beside it as kernel_model_GBps."""
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    row = kernels[dom]
    mean_s = row["mean_launch_ms"] * 1e-3
    alg = ba / 2 * ncell
    ach = alg / mean_s / 1e9 if mean_s > 0 else 0.0
    return {"kernel": dom, "bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS,
            "unit": "GB/s", "frac": ach / PEAK_HBM_GBS,
            "traffic": measured_traffic(dom, workload),
            "algorithmic_bytes_per_launch": alg,
            "algorithmic_bytes_rule": f"B_alg/2 = {ba / 2:g} B per real cell per substep x "
                                      f"{ncell} cells",
            "mean_launch_ms": row["mean_launch_ms"],
            "kernel_model_bytes_per_launch": row["model_bytes_per_launch"],
            "kernel_model_GBps": row["model_GBps"],
            "share_of_step": row["ms_per_step"] / sum(v["ms_per_step"] for v in kernels.values())}


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch of a kernel from the committed rocprofv3 counter
    summary (profiles/traffic_<workload>.json), or None."""
    path = os.path.join(REPO, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    row = table.get("kernels", {}).get(kernel)
    return row.get("hbm_bytes_per_launch") if row else None


def c2p_health(lib, ctx, ntot):
    """The last cons2prim's unconverged share (solves that did not meet the
    stopping rule within 30 Newton iterations, recorded 63 by dh_c2p_info)
    and whether the state is finite (the report's total mass, a sum over
    every real cell, is finite)."""
    import ctypes as C
    import numpy as np
    info = np.zeros(ntot, dtype=np.uint8)
    if lib.dh_c2p_info(ctx, info.ctypes.data_as(C.c_void_p)):
        raise RuntimeError(lib.dh_last_error().decode())
    solved = int((info != 0).sum())
    rep = (C.c_double * 16)()
    if lib.dh_report(ctx, None, 0, rep):
        raise RuntimeError(lib.dh_last_error().decode())
    return {"unconverged_fraction": float(((info & 63) == 63).sum()) / max(solved, 1),
            "state_finite": bool(math.isfinite(rep[7]))}


def ensemble_members(w):
    """(Num_R, Mass_Ratio) of every member: Nr uniform in {64, 128, 256, 512}
    and q log-uniform in [1e-5, 1e-3], from the workload's seed."""
    import numpy as np
    rng = np.random.default_rng(w["seed"])
    return [(int(rng.choice([64, 128, 256, 512])), float(10.0 ** rng.uniform(-5.0, -3.0)))
            for _ in range(w["members"])]


def run_ensemble(args, w, cfg):
    """Independent runs stepped together.  Packed (the default): one
    ensemble (dh_ens_*) whose timestep launches each kernel once for all
    members, with one CFL reduction over all of them and one copy of their
    dts per step, recorded as one HIP graph.  Streams: each member its own
    context, stream and recorded step, its dt fetched one by one.
    value = all members' zone-updates per second."""
    import ctypes as C
    from disco_amd import capi
    from disco_amd.host_lib import Ensemble, HostRun
    steps = args.steps if args.steps is not None else w["steps"]
    warmup = args.warmup if args.warmup is not None else w["warmup"]
    members = ensemble_members(w)
    packed = args.ensemble_mode == "packed"
    tmp = tempfile.mkdtemp(prefix="bench_ens_")
    t0 = time.perf_counter()
    runs = []
    for i, (nr, q) in enumerate(members):
        par = cfg.par_dict(Num_R=nr, Mass_Ratio=q, Autotune=0 if packed else args.autotune)
        runs.append(HostRun(w["config"], state_io.write_par(os.path.join(tmp, str(i)), par)))
    t_setup = time.perf_counter() - t0
    ncell = sum(r.real_cells for r in runs)
    print(f"bench {args.workload}: set-up of {len(runs)} members, {ncell} cells in "
          f"{t_setup:.1f} s", file=sys.stderr, flush=True)
    lib = capi.lib()
    ens = None
    if packed:
        ens = Ensemble(runs, graphs=args.graphs == "on")
        step_all, sync_all, tune = ens.step, ens.sync, 0
    else:
        for r in runs:
            if lib.dh_enable_graphs(C.c_void_p(r.ctx), 1 if args.graphs == "on" else 0):
                raise RuntimeError(lib.dh_last_error().decode())

        def step_all():
            dts = [r.getmindt() for r in runs]
            for r, dt in zip(runs, dts):
                r.timestep(dt)

        def sync_all():
            for r in runs:
                if lib.dh_sync(C.c_void_p(r.ctx)):
                    raise RuntimeError(lib.dh_last_error().decode())
        tune = args.autotune

    for _ in range(tune):  # the members' launch tuning (set-up, untimed)
        step_all()
    for _ in range(warmup):
        step_all()
    sync_all()
    t0 = time.perf_counter()
    for _ in range(steps):
        step_all()
    sync_all()
    secs = time.perf_counter() - t0
    print(f"bench {args.workload}: {steps} timed steps in {secs:.2f} s", file=sys.stderr, flush=True)
    if ens is not None:
        ens.close()
    for r in runs:
        r.close()
    value = ncell * steps / secs
    counts = {n: sum(1 for m in members if m[0] == n) for n in (64, 128, 256, 512)}
    ba = b_alg(cfg.nq, cfg.nf)
    out = {
        "metric": "zone-updates/s on 1 MI355X (%HBM roofline)",
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": secs / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (initial conditions from the workload's Initial plugin)",
        "config": {"workload": w["desc"], "name": args.workload, "graphs": args.graphs == "on",
                   "ensemble": "one launch per kernel for all members (dh_ens)" if packed
                   else "one context and stream per member",
                   "members": len(members), "members_by_Nr": counts, "real_cells": ncell,
                   "setup_s": t_setup},
        "roofline": None,
        "step_roofline": {"b_alg_bytes_per_zone_update": ba, "achieved": ba * value / 1e9,
                          "peak": PEAK_HBM_GBS, "unit": "GB/s",
                          "frac": ba * value / 1e9 / PEAK_HBM_GBS},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg, w["cpu_over"], w["cpu_steps"])
    return out


# The GRMHD workload of the cons2prim comparison: its first `harvest` steps
# (plain guess) give the training data, then `steps` are timed.
C2P_COMPARE = dict(workload="kerr_torus", harvest=50, every=5, rows=20000, epochs=30,
                   steps=10)


def grmhd_c2p_compare(args):
    """The metric's second half: the GRMHD step with the learned cons2prim
    guess against the plain one (the previous primitives), and with the
    better of the two starts per cell by residual, all stopping when both
    Newton residuals are below 1e-10 (C2P_Stop 1).  The model is
    trained online: the plain run's first `harvest` steps give (features,
    targets) pairs of sampled cells every `every` steps, the engine's own
    training kernels fit the MLP on them (dh_c2p_train), and the learned run
    takes the same steps untimed before its timed ones, so both time the
    same stretch of the torus's evolution."""
    import ctypes as C
    import numpy as np
    from disco_amd import c2p_model, capi
    from disco_amd.host_lib import HostRun
    cc = C2P_COMPARE
    w = WORKLOADS[cc["workload"]]
    cfg = configs.get(w["config"])
    base = cfg.par_dict(**w["over"])
    base.update(C2P_Stop=1, Autotune=args.autotune)
    lib = capi.lib()
    rng = np.random.default_rng(314159)
    tmp = tempfile.mkdtemp(prefix="bench_c2p_")

    def timed(run):
        ctx = C.c_void_p(run.ctx)
        t0 = time.perf_counter()
        for _ in range(cc["steps"]):
            run.timestep(run.getmindt())
        if lib.dh_sync(ctx):
            raise RuntimeError(lib.dh_last_error().decode())
        secs = time.perf_counter() - t0
        st = (C.c_longlong * 4)()
        lib.dh_c2p_stats(ctx, st)
        stats = {"iterations_per_cell": st[1] / max(st[0], 1),
                 "learned_fraction": st[2] / max(st[0], 1),
                 "fallback_fraction": st[3] / max(st[0], 1),
                 **c2p_health(lib, ctx, run.ntot)}
        return run.real_cells * cc["steps"] / secs, secs, stats

    run = HostRun(w["config"], state_io.write_par(os.path.join(tmp, "plain"), base))
    ncell = run.real_cells
    xs, ys = [], []
    t0 = time.perf_counter()
    for n in range(cc["harvest"]):
        run.timestep(run.getmindt())
        if (n + 1) % cc["every"] == 0:
            f, t = c2p_model.harvest_ctx(lib, C.c_void_p(run.ctx), run.ntot)
            if len(f) > cc["rows"]:
                pick = rng.choice(len(f), cc["rows"], replace=False)
                f, t = f[pick], t[pick]
            xs.append(f)
            ys.append(t)
    t_harvest = time.perf_counter() - t0
    v_plain, s_plain, st_plain = timed(run)
    run.close()
    X, Y = np.concatenate(xs), np.concatenate(ys)
    t0 = time.perf_counter()
    params, loss = c2p_model.fit(X, Y, epochs=cc["epochs"], batch=2048)
    t_train = time.perf_counter() - t0
    model = os.path.join(tmp, "c2p_online.f32")
    c2p_model.save(params, model)
    print(f"bench c2p: {len(X)} harvested pairs, trained in {t_train:.1f} s (mse {loss:.2e})",
          file=sys.stderr, flush=True)
    runs = {}
    for name, mode in (("learned", 1), ("better", 2)):
        run = HostRun(w["config"], state_io.write_par(os.path.join(tmp, name),
                                                      dict(base, C2P_Guess=mode, C2P_Model=model)))
        for _ in range(cc["harvest"]):
            run.timestep(run.getmindt())
        runs[name] = timed(run)
        run.close()
    v_learned, s_learned, st_learned = runs["learned"]
    v_better, s_better, st_better = runs["better"]
    print(f"bench c2p: plain {v_plain:.3e}, learned {v_learned:.3e}, better of the two "
          f"{v_better:.3e} zone-updates/s", file=sys.stderr, flush=True)
    return {"workload": w["desc"], "name": cc["workload"], "real_cells": ncell,
            "stop_rule": "both residuals below 1e-10 (C2P_Stop 1)",
            "training": {"pairs": int(len(X)), "harvest_steps": cc["harvest"],
                         "harvest_every": cc["every"], "epochs": cc["epochs"],
                         "mse": loss, "train_s": t_train, "harvest_s": t_harvest,
                         "kernels": "dh_c2p_train (float32 forward, backward, Adam on the GPU)"},
            "timed_steps": cc["steps"],
            "plain": {"value": v_plain, "ms_per_step": s_plain / cc["steps"] * 1e3,
                      "c2p": st_plain},
            "learned": {"value": v_learned, "ms_per_step": s_learned / cc["steps"] * 1e3,
                        "c2p": st_learned},
            "better": {"value": v_better, "ms_per_step": s_better / cc["steps"] * 1e3,
                       "c2p": st_better,
                       "guess": "per cell the start with the smaller residual, the model's "
                                "or the previous primitives (C2P_Guess 2)"},
            "learned_over_plain": v_learned / v_plain,
            "better_over_plain": v_better / v_plain}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD, choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kernel-timers", action="store_true")
    ap.add_argument("--c2p-guess", choices=["previous", "learned"], default="previous",
                    help="grmhd2 cons2prim's initial guess (learned: the committed model)")
    ap.add_argument("--c2p-stop", choices=["reference", "residual"], default="reference",
                    help="grmhd2 cons2prim's stopping rule")
    ap.add_argument("--graphs", choices=["on", "off"], default="on",
                    help="record the RK2 timestep once as a HIP graph and replay it")
    ap.add_argument("--autotune", type=int, default=12,
                    help="steps of launch tuning before the warmup (0: untuned launches)")
    ap.add_argument("--ensemble-mode", choices=["packed", "streams"], default="packed",
                    help="disk_ensemble: one launch per kernel for all members, or a "
                         "context and stream per member")
    ap.add_argument("--no-c2p-compare", action="store_true",
                    help="skip the GRMHD plain vs learned cons2prim guess comparison")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(json.dumps({"error": "disco_amd runs one mesh on one GPU; sharding the annulus "
                          "mesh over several GPUs is not implemented", "n_gpus": args.gpus}))
        return 2

    import ctypes as C
    from disco_amd import capi
    from disco_amd.host_lib import HostRun

    w = WORKLOADS[args.workload]
    if "members" in w:
        print(json.dumps(run_ensemble(args, w, configs.get(w["config"]))), flush=True)
        return 0
    steps = args.steps if args.steps is not None else w["steps"]
    warmup = args.warmup if args.warmup is not None else w["warmup"]
    cfg = configs.get(w["config"])
    par = cfg.par_dict(**w["over"])
    gr = cfg.hydro == "grmhd2"
    if gr and args.c2p_guess == "learned":
        from disco_amd import c2p_model
        par.update(C2P_Guess=1, C2P_Model=c2p_model.MODEL_PATH)
    if gr and args.c2p_stop == "residual":
        par["C2P_Stop"] = 1
    tune_steps = args.autotune
    par["Autotune"] = tune_steps
    tmp = tempfile.mkdtemp(prefix="bench_")
    par_path = state_io.write_par(tmp, par)
    t_setup = time.perf_counter()
    run = HostRun(w["config"], par_path)
    t_setup = time.perf_counter() - t_setup
    print(f"bench {args.workload}: set-up of {run.real_cells} cells in {t_setup:.1f} s",
          file=sys.stderr, flush=True)
    lib = capi.lib()
    ctx = C.c_void_p(run.ctx)
    if lib.dh_enable_graphs(ctx, 1 if args.graphs == "on" else 0):
        raise RuntimeError(lib.dh_last_error().decode())

    def sync():
        if lib.dh_sync(ctx):
            raise RuntimeError(lib.dh_last_error().decode())

    # the launch tuner's steps (dh_autotune, in.par Autotune) are part of
    # the set-up: untimed, before the warmup
    t_tune = time.perf_counter()
    for _ in range(tune_steps):
        run.timestep(run.getmindt())
    sync()
    t_tune = time.perf_counter() - t_tune
    tuning = capi.read_tuning(lib, ctx)
    for _ in range(warmup):
        run.timestep(run.getmindt())
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        run.timestep(run.getmindt())
    sync()
    secs = time.perf_counter() - t0
    print(f"bench {args.workload}: {steps} timed steps in {secs:.2f} s", file=sys.stderr, flush=True)
    ncell = run.real_cells
    value = ncell * steps / secs
    nq, nf = cfg.nq, cfg.nf
    mhd = cfg.hydro in ("mhd", "grmhd2")

    nt = C.c_long()
    lib.dh_sizes(ctx, C.byref(nt), None, None)
    ntot = nt.value
    nfaces = []
    for dim in (1, 2):
        n = C.c_long()
        lib.dh_face_count(ctx, dim, C.byref(n))
        nfaces.append(n.value)
    nfr, nfz = nfaces

    # cells of one z ghost layer (Np does not depend on z); the bc_* models
    # cover meshes whose copied sides are the z sides only
    bc = cfg.dh_bc(par)
    ng = 2  # ghost annuli per side (NUM_G)
    nz_tot = int(par.get("Num_Z", 1)) + sum(ng for s in (2, 3) if bc[s] != configs.BC_NONE)
    z_copied = any(bc[s] in (configs.BC_ZEROGRAD, configs.BC_DIODE) for s in (2, 3))
    r_copied = any(bc[s] in (configs.BC_ZEROGRAD, configs.BC_DIODE) for s in (0, 1))
    zlayer = ntot // nz_tot if z_copied and not r_copied else 0

    roofline = None
    kernels = None
    if not args.no_kernel_timers:
        tsteps = max(2, min(steps, 5))
        lib.dh_enable_timers(ctx, 1)
        for _ in range(tsteps):
            run.timestep(run.getmindt())
        sync()
        cap = 64
        names = (C.c_char * 48 * cap)()
        ms = (C.c_double * cap)()
        calls = (C.c_long * cap)()
        n = C.c_int()
        lib.dh_read_timers(ctx, cap, names, ms, calls, C.byref(n))
        lib.dh_enable_timers(ctx, 0)
        kernels = {}
        for k in range(n.value):
            nm = bytes(names[k]).split(b"\0")[0].decode()
            alg = kernel_bytes(nm, nq, mhd, ntot, ncell, nfr, nfz, zlayer)
            mean_s = ms[k] / max(calls[k], 1) * 1e-3
            kernels[nm] = {"ms_per_step": ms[k] / tsteps, "launches_per_step": calls[k] / tsteps,
                           "mean_launch_ms": mean_s * 1e3,
                           "model_bytes_per_launch": alg,
                           "model_GBps": alg / mean_s / 1e9 if mean_s > 0 and alg else None}
        roofline = dominant_roofline(kernels, b_alg(nq, nf), ncell, args.workload)

    ba = b_alg(nq, nf)
    step_ach = ba * value / 1e9
    out = {
        "metric": "zone-updates/s on 1 MI355X (%HBM roofline)",
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": secs / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (initial conditions from the workload's Initial plugin, seeded noise)",
        "config": {"workload": w["desc"], "name": args.workload, "graphs": args.graphs == "on", "real_cells": ncell,
                   "cells_with_ghosts": ntot, "r_faces": nfr, "z_faces": nfz,
                   "setup_s": t_setup, "tuning_steps": tune_steps, "tuning_s": t_tune},
        "tuning": tuning,
        "roofline": roofline,
        "step_roofline": {"b_alg_bytes_per_zone_update": ba, "achieved": step_ach,
                          "peak": PEAK_HBM_GBS, "unit": "GB/s",
                          "frac": step_ach / PEAK_HBM_GBS},
        "kernels": kernels,
        "cpu_baseline": None,
    }
    if gr:
        st = (C.c_longlong * 4)()
        if lib.dh_c2p_stats(ctx, st) == 0 and st[0]:
            out["c2p"] = {"guess": args.c2p_guess, "stop": args.c2p_stop,
                          **c2p_health(lib, ctx, ntot),
                          "iterations_per_cell": st[1] / st[0],
                          "learned_fraction": st[2] / st[0], "fallback_fraction": st[3] / st[0]}
    run.close()
    # the measured streaming bandwidth beside the published 8 TB/s
    cb = copy_bandwidth()
    if cb:
        out["step_roofline"]["measured_copy"] = cb
        out["step_roofline"]["frac_of_measured_copy"] = step_ach / cb["GBps"]
        if roofline:
            roofline["frac_of_measured_copy"] = roofline["achieved"] / cb["GBps"]
    if not args.no_c2p_compare:
        out["grmhd_c2p"] = grmhd_c2p_compare(args)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg, w["cpu_over"], w["cpu_steps"])
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
