#!/usr/bin/env python3
"""Throughput of the Disco timestep on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one full RK2 timestep (getmindt + timestep, both through the
reference-signature host entry points of libdisco_host_<config>.so) over the
whole mesh, with the state resident in HBM.  The timed region is exactly K
steps between device synchronisations.  Prints one JSON line:

  value        zone-updates/s = real cells x K / seconds
  roofline     the dominant kernel (most device time per step): algorithmic
               bytes per launch / its mean launch time (HIP events on the
               library's stream), against the 8 TB/s HBM3E peak
  step_roofline  B_alg x zone-updates/s against 8 TB/s, with
               B_alg = 40 (NQ + NF + 1) bytes per zone-update (the state
               {cons, Phi, piph} read and written over the RK2 step)
  cpu_baseline the reference build (oracle/_ref) on this host, one core, on a
               bounded number of steps of the same mesh
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

from disco_amd import configs, state_io  # noqa: E402

PEAK_HBM_GBS = 8000.0

# name -> (config, in.par overrides, description, cpu-baseline steps)
WORKLOADS = {
    "euler_yee": ("euler_yee", {},
                  "2-D Euler isentropic vortex, Nr=128 polar mesh (51,809 cells), "
                  "HLLC, PLM 1.5, RK2, CFL 0.5, rigid mesh rotation", 150),
}
DEFAULT_WORKLOAD = "euler_yee"


def b_alg(nq: int, nf: int) -> int:
    return 40 * (nq + nf + 1)


def kernel_bytes(name: str, nq: int, ntot: int, nreal: int, nfr: int, nfz: int) -> float:
    """Algorithmic bytes of one launch of a kernel family: every input array
    element the kernel needs read once and every output written once."""
    d = 8
    if name == "update":
        # cons, rkcons (second substep: half the launches), prim in/out,
        # phi/r/z face deltas, geometry in/out
        return d * (nq * nreal * (1 + 0.5 + 1 + 1 + 1) + nq * (nreal + nfr + nfz) + 5 * ntot)
    if name in ("flux_r", "flux_z"):
        nf = nfr if name == "flux_r" else nfz
        return 4 * 2 * nf + d * 2 * nf + d * (3 * nq * ntot + 2 * ntot) + d * nq * nf
    if name in ("face_slopes_r", "face_slopes_z"):
        nf = nfr if name == "face_slopes_r" else nfz
        return d * (2 * nq * ntot + 2 * ntot) + 4 * 2 * nf + d * 2 * nf + d * nq * nf
    if name in ("grad_r", "grad_z"):
        nf = nfr if name == "grad_r" else nfz
        return d * (nq * nf + nf + 2 * ntot + ntot) + d * nq * ntot
    if name == "plm_phi":
        return d * (nq * ntot + ntot + nq * ntot)
    if name == "flux_phi":
        return d * (2 * nq * ntot + 3 * ntot + nq * nreal)
    if name == "wcell":
        return d * 2 * ntot
    if name == "ring_start":
        return d * ntot
    if name == "mindt":
        return d * (nq * nreal + 3 * ntot)
    return 0.0


def cpu_baseline(cfg, par, steps):
    exe = configs.ref_binary(cfg)
    if not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory(prefix="bench_ref_") as d:
        state_io.write_par(d, par)
        r = subprocess.run([exe, d, str(steps)], cwd=d, capture_output=True, text=True,
                           timeout=600)
        if r.returncode != 0:
            return None
        line = [l for l in r.stdout.splitlines() if l.startswith("{")][-1]
        res = json.loads(line)
    return {"value": res["zone_updates_per_s"], "unit": "zone-updates/s", "cores": 1,
            "kind": "reference",
            "sample": f"{steps} RK2 steps of the full mesh ({res['cells']} cells) with the "
                      f"reference timestep built from its sources (gcc -O3, serial), "
                      f"{res['seconds']:.1f} s"}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD, choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kernel-timers", action="store_true")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding of the "
                          "annulus mesh is described in DESIGN.md but not built",
                          "n_gpus": args.gpus}))
        return 2

    import ctypes as C
    from disco_amd import capi
    from disco_amd.host_lib import HostRun

    cname, over, desc, cpu_steps = WORKLOADS[args.workload]
    cfg = configs.get(cname)
    par = cfg.par_dict(**over)
    tmp = tempfile.mkdtemp(prefix="bench_")
    par_path = state_io.write_par(tmp, par)
    run = HostRun(cname, par_path)
    lib = capi.lib()
    ctx = C.c_void_p(run.ctx)

    def sync():
        rc = lib.dh_sync(ctx)
        if rc:
            raise RuntimeError(lib.dh_last_error().decode())

    for _ in range(args.warmup):
        run.timestep(run.getmindt())
    sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run.timestep(run.getmindt())
    sync()
    t1 = time.perf_counter()
    secs = t1 - t0
    ncell = run.real_cells
    value = ncell * args.steps / secs
    nq, nf = cfg.nq, cfg.nf

    roofline = None
    kernels = None
    if not args.no_kernel_timers:
        lib.dh_enable_timers(ctx, 1)
        for _ in range(args.steps):
            run.timestep(run.getmindt())
        sync()
        cap = 64
        names = (C.c_char * 48 * cap)()
        ms = (C.c_double * cap)()
        calls = (C.c_long * cap)()
        n = C.c_int()
        lib.dh_read_timers(ctx, cap, names, ms, calls, C.byref(n))
        lib.dh_enable_timers(ctx, 0)
        nt = C.c_long()
        lib.dh_sizes(ctx, C.byref(nt), None, None)
        ntot = nt.value
        # face counts: every pair of neighbouring annuli has Np[a] + Np[b] faces
        g = state_io  # noqa
        nfr, nfz = _face_counts(par)
        kernels = {}
        for i in range(n.value):
            nm = bytes(names[i]).split(b"\0")[0].decode()
            kernels[nm] = {"ms_per_step": ms[i] / args.steps, "launches": calls[i],
                           "mean_ms": ms[i] / max(calls[i], 1)}
        dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
        alg = kernel_bytes(dom, nq, ntot, ncell, nfr, nfz)
        mean_s = kernels[dom]["mean_ms"] * 1e-3
        ach = alg / mean_s / 1e9 if mean_s > 0 else 0.0
        roofline = {"kernel": dom, "bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS,
                    "unit": "GB/s", "frac": ach / PEAK_HBM_GBS, "traffic": None,
                    "algorithmic_bytes_per_launch": alg,
                    "mean_launch_ms": kernels[dom]["mean_ms"]}

    ba = b_alg(nq, nf)
    step_ach = ba * value / 1e9
    out = {
        "metric": "zone-updates/s on 1 MI355X (%HBM roofline)",
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": secs / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (initial conditions from the workload's Initial plugin)",
        "config": {"workload": desc, "name": args.workload, "real_cells": ncell},
        "roofline": roofline,
        "step_roofline": {"b_alg_bytes_per_zone_update": ba, "achieved": step_ach,
                          "peak": PEAK_HBM_GBS, "unit": "GB/s",
                          "frac": step_ach / PEAK_HBM_GBS},
        "kernels": kernels,
        "cpu_baseline": None,
    }
    run.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg, par, cpu_steps)
    print(json.dumps(out), flush=True)
    return 0


def _face_counts(par):
    """r- and z-face counts of the mesh (sum over neighbouring annulus pairs)."""
    import types
    d = types.SimpleNamespace()
    nr = int(par["Num_R"])
    nz = int(par.get("Num_Z", 1))
    ng = 2
    ngra = 0 if par.get("NoBC_Rmin", 0) else ng
    ngrb = 0 if par.get("NoBC_Rmax", 0) else ng
    ngza = 0 if (nz == 1 or par.get("NoBC_Zmin", 0)) else ng
    ngzb = 0 if (nz == 1 or par.get("NoBC_Zmax", 0)) else ng
    NR, NZ = nr + ngra + ngrb, nz + ngza + ngzb
    rmin, rmax = float(par["R_Min"]), float(par["R_Max"])
    aspect = float(par.get("aspect", 1.0))
    npr = []
    for j in range(NR):
        x1 = (j - ngra + 1) / nr
        x0 = (j - ngra) / nr
        rp = rmin + x1 * (rmax - rmin)
        rm = rmin + x0 * (rmax - rmin)
        npr.append(max(4, int(2 * 3.141592653589793 / (aspect * (rp - rm) / rp))))
    nfr = sum(npr[j] + npr[j + 1] for j in range(NR - 1)) * NZ
    nfz = sum(2 * npr[j] for j in range(NR)) * (NZ - 1)
    del d
    return nfr, nfz


if __name__ == "__main__":
    sys.exit(main())
