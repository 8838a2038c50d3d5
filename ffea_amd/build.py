"""Build ffea_amd/libffea_hip.so (gfx950) from ffea_amd/csrc with hipcc, in-tree.

    python -m ffea_amd.build        # or ffea_amd.build.build()

The FP64 parity path needs strict rounding: host and device code are compiled with
-ffp-contract=off (no FMA contraction).  Packed FP32 instructions are disabled on the
device (-packed-fp32-ops).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libffea_hip.so")
ARCH = os.environ.get("FFH_OFFLOAD_ARCH", "gfx950")


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    hipcc = _hipcc()
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    plan_o = os.path.join(obj_dir, "ffh_plan.o")
    kern_o = os.path.join(obj_dir, "ffh_kernels.o")
    common = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function", "-pthread"]
    cmds = [
        ["g++", *common, "-c", os.path.join(CSRC, "ffh_plan.cpp"), "-o", plan_o],
        [hipcc, f"--offload-arch={ARCH}", *common, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-c", os.path.join(CSRC, "ffh_kernels.hip"), "-o", kern_o],
        [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", OUT + ".tmp", plan_o, kern_o],
    ]
    for c in cmds:
        if verbose:
            print(" ".join(c), flush=True)
        r = subprocess.run(c, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"build failed: {' '.join(c[:2])} ...")
        elif verbose and r.stderr.strip():
            sys.stderr.write(r.stderr)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
