"""Build the HIP library in-tree with hipcc for gfx950.

    python -m kestrel_amd.build

Produces kestrel_amd/libkestrel_amd.so and, where AMD flang is installed, the
Fortran host driver kestrel_amd/fortran/kst_host.  The FP64 path must reproduce the
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
HEADERS = [os.path.join(HERE, "csrc", "kst_physics.hpp"), os.path.join(HERE, "csrc", "kst_morpho.hpp"),
           os.path.join(HERE, "csrc", "kst_output.hpp"), os.path.join(ROOT, "include", "kestrel_step.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
         "-ffp-contract=off", "-fno-fast-math",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FORTRAN = os.path.join(HERE, "fortran")
HOST = os.path.join(FORTRAN, "kst_host")


def build_fortran_host(verbose: bool = False) -> str | None:
    """The Fortran host driver (kst_host.f90 over the kestrel_step_binding
    module), linked to the library with an rpath to it; skipped without flang."""
    if not os.path.exists(FLANG):
        return None
    mod = os.path.join(FORTRAN, "_mod")
    os.makedirs(mod, exist_ok=True)
    objs = []
    for src in ("kestrel_step.f90", "kst_host.f90"):
        obj = os.path.join(mod, src.replace(".f90", ".o"))
        cmd = [FLANG, "-O2", "-J" + mod, "-c", os.path.join(FORTRAN, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    cmd = [FLANG, "-o", HOST + ".tmp", *objs, "-L" + HERE, "-lkestrel_amd", "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(HOST + ".tmp", HOST)
    return HOST


def build(verbose: bool = False) -> str:
    cmd = [HIPCC, *FLAGS, "-o", LIB + ".tmp", *SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    build_fortran_host(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print(LIB)
