"""Summaries of rocprofv3 runs of the timestep, for profiles/.

    python -m disco_amd.profile_summary kernels --db TRACE.db --out FILE.md [--cmd "..."]
    python -m disco_amd.profile_summary traffic --fetch F.db --write W.db \\
        --copy-fetch CF.db --copy-write CW.db --workload NAME --out FILE.json [--cmd "..."]
    python -m disco_amd.profile_summary copy-probe

`kernels` tabulates a --kernel-trace run per kernel (calls, mean and total
duration), with each kernel's name in the engine's timers.  `traffic` turns
two --pmc runs (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch) into HBM bytes
per launch of every kernel, scaled by what the same counters read for a
plain copy of known size (`copy-probe`: one 1 GiB device-to-device copy),
which bench.py reports next to the algorithmic bytes.
"""
from __future__ import annotations

import argparse
import json
import re
import sqlite3
import sys
from collections import defaultdict

# kernel symbol (as rocprofv3 prints it) -> the engine's timer name
_NAMES = [
    (r"KBoth<dh::KWcell<", "wcell_plm_phi"), (r"KBoth<dh::KFluxPhi<", "flux_phi_r"),
    (r"KPlmPhi<", "plm_phi"),
    (r"k_ring_start", "ring_start"),
    (r"KCellRanges<1>", "ranges_r"), (r"KCellRanges<2>", "ranges_z"),
    (r"KFaces<1>", "faces_r"), (r"KFaces<2>", "faces_z"),
    (r"KGrad<\d+, 1>", "grad_r"), (r"KGrad<\d+, 2>", "grad_z"),
    (r"KFluxPhi<", "flux_phi"),
    (r"KFluxTrans<.*, 1>", "flux_r"), (r"KFluxTrans<.*, 2>", "flux_z"),
    (r"KEdgesCollect", "edges_collect"),
    (r"KEdgesFinal", "edges_final"),
    (r"KUpdateB", "update_B"),
    (r"KLineSteps<1>", "line_steps_r"), (r"KLineSteps<2>", "line_steps_z"),
    (r"KEdgeSegments", "edge_segments"), (r"KEdgeAdjust", "edge_adjust"),
    (r"KUpdate<", "update"),
    (r"KFlipFlags<1>", "flip_flags_r"), (r"KFlipFlags<2>", "flip_flags_z"),
    (r"k_flips", "flips"),
    (r"KBfacesC2p<", "bfaces_c2p"),
    (r"KBcScan", "bc_scan"), (r"KBcFill<", "bc_fill"),
    (r"k_mindt<", "mindt"), (r"k_bc_scan", "bc_scan"), (r"k_diag<", "diagnostics"),
    (r"KDiagAvg", "diag_avg"),
    (r"KWcell<", "wcell"), (r"KWcellAvg", "wcell"),
    (r"KInitCells<", "init"),
]


def timer_name(kernel: str) -> str:
    head = kernel.split("(")[0] if kernel.startswith("k_") else kernel
    for pat, name in _NAMES:
        if re.search(pat, head):
            return name
    return ""


def _rows(db: str, sql: str):
    con = sqlite3.connect(db)
    try:
        return con.execute(sql).fetchall()
    finally:
        con.close()


def functor_of(kernel: str) -> str:
    """The functor a k_run / k_run_xcd launch runs, whatever its launch
    variant (storage order, XCD-contiguous, or wrapped in a group order);
    other kernels are their own."""
    m = re.match(r"void k_run(?:_xcd)?<(.*)>\((.*), long\)$", kernel)
    if not m:
        return kernel
    f = m.group(2)
    g = re.match(r"dh::Grouped<(.*)>$", f)
    return g.group(1).strip() if g else f.strip()


def after_tuning(rows):
    """The rows (kernel, dispatch_id, value) of each functor's last launch
    variant from the dispatch after the last one of any other variant of the
    same functor: the launches that follow the launch tuner's sweep."""
    last = defaultdict(lambda: -1)
    for name, did, _ in rows:
        last[name] = max(last[name], did)
    by_functor = defaultdict(list)
    for name in last:
        by_functor[functor_of(name)].append(name)
    keep = {}
    for names in by_functor.values():
        final = max(names, key=lambda n: last[n])
        keep[final] = max([last[n] for n in names if n != final], default=-1)
    return [r for r in rows if r[0] in keep and r[1] > keep[r[0]]]


def kernel_stats(db: str, tuned: bool = False):
    """{kernel: (calls, total_ns, mean_ns)} of a --kernel-trace database
    (tuned: only the launches after the tuner's sweep)."""
    rows = _rows(db, "select name, dispatch_id, duration from kernels")
    if tuned:
        rows = after_tuning(rows)
    acc = defaultdict(list)
    for name, _, dur in rows:
        acc[name].append(dur)
    return {k: (len(v), sum(v), sum(v) / len(v)) for k, v in acc.items()}


def counter_means(db: str, counter: str, tuned: bool = False):
    """{kernel: (dispatches, mean value)} of one --pmc counter (tuned: only
    the launches after the tuner's sweep)."""
    rows = _rows(db, "select kernel_name, dispatch_id, value from counters_collection "
                     f"where counter_name = '{counter}'")
    if tuned:
        rows = after_tuning(rows)
    acc = defaultdict(list)
    for name, _, value in rows:
        acc[name].append(value)
    return {k: (len(v), sum(v) / len(v)) for k, v in acc.items()}


def _kernel_table(stats):
    total = sum(v[1] for v in stats.values())
    lines = ["| kernel | timer name | calls | mean (ms) | total (ms) | share |",
             "|---|---|---|---|---|---|"]
    for name, (calls, tot, mean) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        short = name if len(name) < 90 else name[:87] + "..."
        lines.append(f"| `{short}` | {timer_name(name)} | {calls} | {mean / 1e6:.4f} | "
                     f"{tot / 1e6:.3f} | {100 * tot / total:.1f}% |")
    return lines


def cmd_kernels(args):
    lines = [f"# Kernel trace: {args.title}", ""]
    if args.cmd:
        lines += ["Command:", "", "```", args.cmd, "```", ""]
    lines += ["## After the launch tuner's sweep", "",
              "Each kernel in the launch variant the tuner kept, from its first launch after "
              "the last launch of any other variant (the warmup and timed steps).", ""]
    lines += _kernel_table(kernel_stats(args.db, tuned=True))
    lines += ["", "## Every dispatch of the run", ""]
    lines += _kernel_table(kernel_stats(args.db))
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {args.out}")


def _copy_factor(fetch_db: str, write_db: str):
    """(read factor, write factor): true bytes / counter bytes of the copy probe."""
    nbytes = 1 << 30
    f = counter_means(fetch_db, "FETCH_SIZE")
    w = counter_means(write_db, "WRITE_SIZE")
    # the probe's one large copy is the dispatch with the largest counts
    fk = max(f, key=lambda k: f[k][1])
    wk = max(w, key=lambda k: w[k][1])
    return nbytes / (f[fk][1] * 1024), nbytes / (w[wk][1] * 1024), fk


def cmd_traffic(args):
    rf, wf, probe = _copy_factor(args.copy_fetch, args.copy_write)
    fetch = counter_means(args.fetch, "FETCH_SIZE", tuned=args.tuned)
    write = counter_means(args.write, "WRITE_SIZE", tuned=args.tuned)
    kernels = {}
    for name in sorted(set(fetch) | set(write)):
        tn = timer_name(name)
        if not tn:
            continue
        nf, fv = fetch.get(name, (0, 0.0))
        nw, wv = write.get(name, (0, 0.0))
        rd = fv * 1024 * rf
        wr = wv * 1024 * wf
        row = kernels.setdefault(tn, {"kernel": name, "dispatches": max(nf, nw),
                                      "read_bytes_per_launch": 0.0,
                                      "write_bytes_per_launch": 0.0})
        row["read_bytes_per_launch"] += rd
        row["write_bytes_per_launch"] += wr
    for row in kernels.values():
        row["hbm_bytes_per_launch"] = row["read_bytes_per_launch"] + row["write_bytes_per_launch"]
    out = {"workload": args.workload,
           "counters": "FETCH_SIZE and WRITE_SIZE (KiB per dispatch), one rocprofv3 --pmc run each",
           "launches": ("each kernel in the launch variant the tuner kept, after its sweep"
                        if args.tuned else "every dispatch"),
           "scale": {"read": rf, "write": wf,
                     "probe": f"1 GiB device-to-device copy ({probe.split('(')[0]})"},
           "command": args.cmd, "kernels": kernels}
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {args.out}: read x{rf:.3f}, write x{wf:.3f}")


def cmd_copy_probe(args):
    import torch
    n = 1 << 27  # 1 GiB of doubles
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    print("copy probe: 1 GiB device-to-device")


FP64_COUNTERS = ("SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_FMA_F64",
                 "SQ_INSTS_VALU_TRANS_F64")
PEAK_FP64_TFLOPS = 78.6  # MI355X vector FP64


def cmd_fp64(args):
    """FP64 rate of every kernel: wavefront FP64 instructions per dispatch
    (one --pmc run with FP64_COUNTERS) times 64 lanes, an FMA counting two
    operations, over the kernel's mean duration in a --kernel-trace run of
    the same workload (the launches after the tuner's sweep, in the variant
    it kept; every launch variant of a kernel runs the same instructions)."""
    stats = kernel_stats(args.trace, tuned=True)
    # instruction counts per functor: every launch variant runs the same code
    per = {}
    for c in FP64_COUNTERS:
        acc = defaultdict(list)
        for name, (n, mean) in counter_means(args.pmc, c).items():
            acc[functor_of(name)] += [mean] * n
        per[c] = {f: sum(v) / len(v) for f, v in acc.items()}
    rows = []
    for name, (calls, tot, mean_ns) in stats.items():
        fn = functor_of(name)
        if fn not in per[FP64_COUNTERS[0]]:
            continue
        add, mul, fma, trans = (per[c].get(fn, 0.0) for c in FP64_COUNTERS)
        flops = 64.0 * (add + mul + trans + 2.0 * fma)
        tf = flops / mean_ns / 1e3
        rows.append((name, mean_ns / 1e6, flops, tf, tf / PEAK_FP64_TFLOPS))
    rows.sort(key=lambda r: -r[1])
    lines = [f"# FP64 rate: {args.title}", "",
             f"FP64 operations per dispatch = 64 lanes x (ADD + MUL + TRANS + 2 FMA) wavefront "
             f"instructions (rocprofv3 counters {', '.join(FP64_COUNTERS)}); rate over the mean "
             f"duration of the same kernel in a kernel-trace run; peak {PEAK_FP64_TFLOPS} TFLOP/s "
             f"(MI355X vector FP64).", ""]
    if args.cmd:
        lines += ["Commands:", "", "```", args.cmd, "```", ""]
    lines += ["| kernel | timer name | mean (ms) | FP64 ops per dispatch | TFLOP/s | of peak |",
              "|---|---|---|---|---|---|"]
    for name, ms, flops, tf, frac in rows:
        short = name if len(name) < 90 else name[:87] + "..."
        lines.append(f"| `{short}` | {timer_name(name)} | {ms:.4f} | {flops:.3e} | {tf:.2f} | "
                     f"{100 * frac:.1f}% |")
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {args.out}")


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("--db", required=True)
    k.add_argument("--out", required=True)
    k.add_argument("--title", default="")
    k.add_argument("--cmd", default="")
    t = sub.add_parser("traffic")
    for a in ("--fetch", "--write", "--copy-fetch", "--copy-write", "--workload", "--out"):
        t.add_argument(a, required=True)
    t.add_argument("--cmd", default="")
    t.add_argument("--tuned", action="store_true",
                   help="only the launches after the launch tuner's sweep")
    sub.add_parser("copy-probe")
    f = sub.add_parser("fp64")
    for a in ("--pmc", "--trace", "--out"):
        f.add_argument(a, required=True)
    f.add_argument("--title", default="")
    f.add_argument("--cmd", default="")
    args = ap.parse_args(argv)
    {"kernels": cmd_kernels, "traffic": cmd_traffic, "copy-probe": cmd_copy_probe,
     "fp64": cmd_fp64}[args.what](args)
    return 0


if __name__ == "__main__":
    sys.exit(main())
