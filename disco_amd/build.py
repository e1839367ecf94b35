"""Builds the native parts in-tree.

    python -m disco_amd.build            # libdisco_hip.so (+ host library)
    python -m disco_amd.build --all      # also the oracle and, when the
                                         # reference sources are mounted,
                                         # oracle/_ref

libdisco_hip.so is compiled for gfx950 only, with packed-FP32 instructions
disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(REPO, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
REFERENCE = "/root/reference"

# -freciprocal-math, for the device code only (the library's host code keeps
# IEEE division): divisions by one divisor may share its reciprocal
# (one correctly rounded 1/y, then multiplications, instead of a full
# division sequence per quotient); the Riemann solvers divide many
# quantities by the same density, enthalpy or wave-speed differences, and
# the flux kernels run 8-12% faster.  A quotient then differs from the
# correctly rounded one by an ulp or so, within the 1e-10 the GPU parity
# tests hold.  The CPU build of the same kernels (tests/cpu_engine) keeps
# IEEE division and no FMA contraction, bit-exact with the reference.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Xarch_device", "-freciprocal-math",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, **kw)


def build_hip(extra=()) -> str:
    out = os.path.join(HERE, "libdisco_hip.so")
    tmp = out + ".tmp"
    srcs = [os.path.join(CSRC, "dh_engine.hip")]
    _run([HIPCC, *HIP_FLAGS, *extra, f"-I{INCLUDE}", f"-I{CSRC}", *srcs, "-o", tmp])
    os.replace(tmp, out)
    return out


# Module selections that have host plugins in disco_amd/plugins: one host
# library and one `disco` driver per selection, as the reference links one
# binary per Makefile_opt.in.
HOST_CONFIGS = {
    "euler_yee": dict(hydro="euler", initial="yee_vortex", planets="none", boundary="polar",
                      analysis="diag_euler", num_c=5, num_n=0, ct_mode=0),
    "euler_yee_n1": dict(hydro="euler", initial="yee_vortex", planets="none", boundary="polar",
                         analysis="diag_disk2d", num_c=5, num_n=1, ct_mode=0),
    "mhd_kep": dict(hydro="mhd", initial="kep_mhd_disk", planets="central", boundary="fixed_r_outflow_z",
                    analysis="diag_mhdmri", num_c=8, num_n=0, ct_mode=2),
    # the MHD disk with a companion (Mass_Ratio > 0) and the Euler diagnostics,
    # which read planet 1
    "mhd_kep_p2": dict(hydro="mhd", initial="kep_mhd_disk", planets="central",
                       boundary="fixed_r_outflow_z", analysis="diag_euler",
                       num_c=8, num_n=0, ct_mode=2),
    "gr_disk": dict(hydro="grmhd2", initial="gr_disk", planets="none",
                    boundary="fixed_r_outflow_z", analysis="diag_euler",
                    num_c=8, num_n=0, ct_mode=2),
    "disk_planet": dict(hydro="euler", initial="planet_disk", planets="planet_circ",
                        boundary="fixed", analysis="diag_euler", num_c=5, num_n=0, ct_mode=0),
    "gr_disk3d": dict(hydro="grmhd2", metric="schw_ks_rpz", initial="gr_disk", planets="none",
                      boundary="fixed_r_outflow_z", analysis="diag_euler",
                      num_c=8, num_n=0, ct_mode=2),
    # the Fishbone-Moncrief torus around a spinning black hole (Kerr-Schild)
    "kerr_torus": dict(hydro="grmhd2", metric="kerr_ks_rpz", initial="fm_torus", planets="none",
                       boundary="fixed_r_outflow_z", analysis="diag_euler",
                       num_c=8, num_n=0, ct_mode=2),
}


def host_defines(c: dict) -> list:
    """The -D flags of one module selection (the reference's Makefile_opt.in
    variables and paul.h sizes)."""
    defs = [f"-DNUM_C={c['num_c']}", f"-DNUM_N={c['num_n']}", f"-DCT_MODE={c['ct_mode']}",
            f'-DBOUNDARY="{c["boundary"]}"', f'-DINITIAL="{c["initial"]}"',
            f'-DANALYSIS="{c["analysis"]}"', f'-DHYDRO="{c["hydro"]}"']
    if "metric" in c:
        defs.append(f'-DMETRIC="{c["metric"]}"')
    return defs


def build_host(configs=None) -> None:
    """libdisco_host_<cfg>.so and disco_<cfg> (host C + plugins, linked to
    libdisco_hip.so)."""
    host = os.path.join(HERE, "host")
    plug = os.path.join(HERE, "plugins")
    for name, c in HOST_CONFIGS.items():
        if configs and name not in configs:
            continue
        defs = host_defines(c)
        srcs = [os.path.join(host, "disco_host.c"),
                os.path.join(plug, "Initial", c["initial"] + ".c"),
                os.path.join(plug, "Planets", c["planets"] + ".c")]
        common = ["gcc", "-O2", "-fPIC", "-Wall", "-Wno-unused-parameter", *defs,
                  f"-I{os.path.join(INCLUDE, 'disco')}", f"-I{INCLUDE}", f"-I{host}", f"-I{CSRC}"]
        link = [f"-L{HERE}", "-ldisco_hip", "-lm", "-Wl,-rpath,$ORIGIN/.."]
        _run(common + ["-shared", *srcs, "-o", os.path.join(host, f"libdisco_host_{name}.so"),
                       *link])
        _run(common + [os.path.join(host, "main.c"), *srcs, "-o",
                       os.path.join(host, f"disco_{name}"), *link])


def build_oracle() -> None:
    _run(["make", "-s", "-f", os.path.join(REPO, "oracle", "Makefile")])


def build_ref() -> bool:
    if not os.path.isdir(REFERENCE):
        print("reference sources not mounted: oracle/_ref not rebuilt", flush=True)
        return False
    _run(["make", "-s", "-j4", "-f", os.path.join(REPO, "oracle", "ref", "Makefile"),
          f"REF={REFERENCE}"])
    return True


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--all", action="store_true")
    args = ap.parse_args(argv)
    build_hip()
    build_host()
    if args.all:
        build_oracle()
        build_ref()


if __name__ == "__main__":
    main()
