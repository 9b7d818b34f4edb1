"""Build the HIP library eqdyna_amd/libeqd.so for gfx950 with hipcc.

Device code is compiled without FMA contraction (-ffp-contract=off) so the
FP64 path performs the reference's rounded operations, and with packed FP32
instructions disabled (-packed-fp32-ops).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libeqd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("EQD_OFFLOAD_ARCH", "gfx950")

SOURCES = ["eqd_kernels.hip", "eqd_mlp.hip", "eqd_precompute.hip", "eqd_api.cpp", "eqd_engine.cpp"]
FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-fopenmp", "-Wno-unused-command-line-argument"]


def build(verbose: bool = False, out: str = LIB, defines=()) -> str:
    objs = []
    tag = "".join(d.replace("=", "") for d in defines)
    for src in SOURCES:
        obj = os.path.join(CSRC, "_build", src + tag + ".o")
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        cmd = [HIPCC, *FLAGS, *[f"-D{d}" for d in defines], "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-fopenmp", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
