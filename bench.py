"""Throughput of the MI355X time step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu]

One "step" is one accepted outer step of IntegrateTo (src/TimeStepper.f90:151-275)
over the active tiles of the workload; the unit of work is the cell-update
(one active cell advanced through one accepted step).  The timed region holds
exactly K steps with the state resident in HBM; the host only enqueues the
step kernels and reads the device's step counter back.

Printed: one JSON line with value = cell-updates/s, the roofline of the
dominant kernel (the stage-2 kernel, timed with HIP events on the library's
stream) and the CPU baseline (the oracle, a C restatement, on a bounded sample
of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of the stage-2 pass per cell (u1 4 + u^n 4 + b0 1 words
# read, u2 4 written; FP64) -- the survey's stage-pass model.
P2_WORDS = 13
STEP_WORDS = 42


def cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor()


def cpu_baseline(wl_name: str, budget_s: float = 20.0) -> dict:
    """The oracle (C restatement, 1 thread) on a bounded sample of the
    workload: the first steps from the same initial state, until budget_s."""
    from kestrel_amd import workloads
    from oracle.oracle import Oracle
    wl = getattr(workloads, wl_name)()
    o = Oracle(wl.cfg)
    o.set_topography(wl.b0v)
    o.set_state(wl.init)
    t0 = time.perf_counter()
    cells = 0
    steps = 0
    while True:
        s = o.step_n(1)
        cells += s["cell_updates"]
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 1000:
            break
    o.close()
    return {"value": cells / el, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"{wl_name}: first {steps} steps ({cells} cell-updates, {el:.1f} s) "
                      f"on {cpu_model()}, 1 thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="plain_cap_2048")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is designed, not built)"}))
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)
    from kestrel_amd import workloads
    from kestrel_amd.runtime import Stepper

    wl = getattr(workloads, args.workload)()
    g = Stepper(wl.cfg)
    g.set_topography(wl.b0v)
    g.set_state(wl.init)
    if args.warmup:
        g.step_n(args.warmup)
    torch.cuda.synchronize()
    # timed region: K steps; every stage kernel is bracketed by HIP events on
    # the library's stream (direct launches)
    g.profile(True)
    t0 = time.perf_counter()
    st = g.step_n(args.steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    stage_ms, launches = g.profile_read()
    g.profile(False)
    # the same K steps again without events, as the library runs them (hipGraph)
    t1 = time.perf_counter()
    st2 = g.step_n(args.steps)
    torch.cuda.synchronize()
    el2 = time.perf_counter() - t1

    cells = st["cell_updates"]
    value = cells / el
    cells_per_step = cells / max(1, st["steps"])
    p2_s = stage_ms[1] * 1e-3
    achieved = cells_per_step * P2_WORDS * 8 / p2_s / 1e9 if p2_s > 0 else 0.0
    res = {
        "metric": "cell-updates/s vs HBM roofline (FP64 plain vs batched); p50/p99 scenario latency",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl.name, "grid": f"{wl.cfg.NX}x{wl.cfg.NY}", "dx_m": wl.cfg.dx,
                   "tiles": f"{wl.cfg.nx_tiles}x{wl.cfg.ny_tiles} of {wl.cfg.nx_per_tile}x{wl.cfg.ny_per_tile}",
                   "active_cells_per_step": cells_per_step, "precision": "fp64 plain",
                   "rhs_evals": st["rhs_evals"], "rejected": st["rejected"],
                   "stage_kernel_ms": stage_ms},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "stage_kernel<2> (RK stage 2)",
                     "algorithmic_bytes_per_cell": P2_WORDS * 8},
        "step_roofline_frac": value * STEP_WORDS * 8 / (HBM_PEAK_GBS * 1e9),
        "graph_run": {"value": st2["cell_updates"] / el2, "ms_per_step": el2 * 1e3 / args.steps,
                      "note": "next K steps, hipGraph launches, no events"},
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline("plain_cap_256_sample", args.cpu_budget)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
