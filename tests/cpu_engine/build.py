"""Builds tests/cpu_engine/_build/libdh_cpu.so: the engine's kernel bodies
compiled as host code (hipcc, host pass only, no FMA contraction)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build", "libdh_cpu.so")
SRC = os.path.join(HERE, "dh_cpu.cpp")
DEPS = [SRC] + [os.path.join(REPO, "disco_amd", "csrc", f) for f in
                ("dh_kernels.h", "dh_pipeline.h", "dh_ct.h", "dh_mesh.h", "dh_geom.h",
                 "dh_euler.h", "dh_mhd.h")] + [os.path.join(REPO, "include", "disco_hip.h")]


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
