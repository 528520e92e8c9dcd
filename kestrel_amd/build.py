"""Build the HIP library in-tree with hipcc for gfx950.

    python -m kestrel_amd.build

Produces kestrel_amd/libkestrel_amd.so.  The FP64 path must reproduce the
reference's arithmetic, so contraction of a*b+c into FMA is disabled, and
packed FP32 instructions are switched off.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libkestrel_amd.so")
SOURCES = [os.path.join(HERE, "csrc", "kst_step.hip")]
HEADERS = [os.path.join(HERE, "csrc", "kst_physics.hpp"), os.path.join(ROOT, "include", "kestrel_step.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
         "-ffp-contract=off", "-fno-fast-math",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def build(verbose: bool = False) -> str:
    cmd = [HIPCC, *FLAGS, "-o", LIB + ".tmp", *SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print(LIB)
