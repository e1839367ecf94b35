"""Builds tests/cpu_engine/_build/libdh_cpu.so: the engine's kernel bodies
compiled as host code (hipcc, host pass only, no FMA contraction)."""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build", "libdh_cpu.so")
SRC = os.path.join(HERE, "dh_cpu.cpp")
DEPS = ([SRC] + sorted(glob.glob(os.path.join(REPO, "disco_amd", "csrc", "*.h")))
        + [os.path.join(REPO, "include", "disco_hip.h")])


def build(force: bool = False) -> str:
    if not force and os.path.exists(OUT) and all(
            os.path.getmtime(OUT) >= os.path.getmtime(d) for d in DEPS):
        return OUT
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    cmd = [hipcc, "-x", "hip", "--cuda-host-only", "-std=c++17", "-O2", "-ffp-contract=off",
           "-fPIC", "-shared", f"-I{os.path.join(REPO, 'include')}",
           f"-I{os.path.join(REPO, 'disco_amd', 'csrc')}", SRC, "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))


def build_driver(config: str) -> str:
    """The `disco` driver of one module selection (disco_amd/host + plugins)
    linked to the host build of the engine instead of libdisco_hip.so."""
    sys.path.insert(0, REPO)
    from disco_amd.build import HOST_CONFIGS, host_defines
    lib = build()
    c = HOST_CONFIGS[config]
    out = os.path.join(HERE, "_build", f"disco_{config}")
    host = os.path.join(REPO, "disco_amd", "host")
    plug = os.path.join(REPO, "disco_amd", "plugins")
    srcs = [os.path.join(host, "main.c"), os.path.join(host, "disco_host.c"),
            os.path.join(plug, "Initial", c["initial"] + ".c"),
            os.path.join(plug, "Planets", c["planets"] + ".c")]
    deps = srcs + [lib, os.path.join(host, "disco_host.h"), os.path.join(REPO, "disco_amd", "build.py")]
    if os.path.exists(out) and all(os.path.getmtime(out) >= os.path.getmtime(s) for s in deps):
        return out
    cmd = ["gcc", "-O2", "-Wall", "-Wno-unused-parameter", *host_defines(c),
           f"-I{os.path.join(REPO, 'include', 'disco')}",
           f"-I{os.path.join(REPO, 'include')}", f"-I{host}",
           f"-I{os.path.join(REPO, 'disco_amd', 'csrc')}", *srcs, "-o", out + ".tmp",
           f"-L{os.path.dirname(lib)}", "-ldh_cpu", "-lm",
           f"-Wl,-rpath,{os.path.dirname(lib)}"]
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out
