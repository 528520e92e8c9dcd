"""Throughput of the MI355X time step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu] [--no-plain]

One "step" is one accepted outer step of IntegrateTo (src/TimeStepper.f90:151-275)
over the active tiles of the workload -- with morphodynamics the Strang step
H(dt) M(2dt) H(dt) -- and the unit of work is the cell-update (one active cell
advanced through one accepted step).

The default workload is the full morphodynamic run: 1024x1024 cells of 5 m,
32 seeded Gaussian hills on a 5 degree ramp, a one-cell flux source of
q*dx = 100 m^3/s with psi = 0.3, Pouliquen drag, Mixed erosion,
Spearman-Manning deposition, eddy viscosity 1e-2, FP64, 2000 steps from the
initial state.  W warm-up steps run on a separate handle; the K timed steps
start from the initial state with everything resident in HBM, launched as the
library runs them (hipGraph), bracketed by device synchronisation.

Printed: one JSON line with value = cell-updates/s of those K steps; the
per-kernel launch times of a second, profiled pass over the same K steps (HIP
events on the library's stream around every launch, launches that did work
only); the roofline of the dominant kernel (the most total time); the plain
FP64 hydraulic path at 2048x2048 with its own roofline; the wall time of one
kst_volume call (CalculateVolume on the device) on that grid; and the CPU
baselines: the reference program itself (its Fortran sources built with
release optimisation, serial, oracle/_ref/time) on a bounded sample of the
same deck -- the first 8 s of simulated time of the morphodynamic run, two
steps of the 2048 x 2048 plain run -- beside the oracle (the C restatement,
1 thread) and, for all host cores, one reference process per core on the
256 x 256 cap release.  Launch shapes come from the library's autotuner
(timed trials per active-set bucket, made in the warm-up).
--volume-only runs the kst_volume leg alone.
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
# Algorithmic FP64 words per cell of one launch, stage-pass model: P1 reads
# u^n 4 + b0 1, writes u^1 4; P2 and P3 read their stage 4 + u^n 4 + b0 1,
# write 4; P4 reads u^3 4 + b0 1, writes momentum 2.  Over a changing bed each
# pass also reads the cell bt (1).  A morphodynamic SSP stage reads u_k 4,
# b0 1, bt_k 1, bt_0 1, w and Hnpsi of u^0 2 and writes bt, w, Hnpsi 3.
STAGE_WORDS = {"hydraulic": [9, 13, 13, 7], "morpho": [10, 14, 14, 8]}
MORPHO_STAGE_WORDS = 12
# One RedistributeCell: 13 words read for each of the 9 cells of its 3x3
# block (4 vertex bt, 4 vertex b0, bt, gamma, w, Hnpsi, b0 of intermed0), 13
# for the corrected cell itself, 6 written per cell of the block and up to 4
# vertices: 188 words per corrected cell.
REDIST_WORDS = 188
# Whole step per cell-update: hydraulic 42 words; Strang step 134 words.
STEP_WORDS = {"hydraulic": 42, "morpho": 134}
# HBM traffic of the dominant kernel from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE runs of this script (mean KiB per productive dispatch, active
# cells per launch of that run), scaled by the counters' reading for a plain
# copy (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0).  Summaries and commands:
# profiles/stage_kernel_plain_cap_2048.md, profiles/kernels_morpho_hills_1024.md.
PMC_FETCH_SCALE, PMC_WRITE_SCALE = 2.0, 1.0
PMC_TRAFFIC = {
    ("plain_cap_2048", "stage_kernel<2>"): (212956.2, 51263.5, 1634048.0),
    ("plain_cap_2048", "stage_kernel<3>"): (202253.5, 51263.5, 1634048.0),
    ("plain_cap_2048", "stage_kernel<4>"): (214820.1, 113545.8, 1634048.0),
    ("morpho_hills_1024", "redistribute_kernel"): (92.3, 35.6, 5795.84),
    ("morpho_hills_1024", "stage_kernel<2>"): (1111.6, 181.9, 5795.84),
    ("morpho_hills_1024", "morpho_kernel<3>"): (645.8, 317.7, 5795.84),
}
# VALU wave-instructions per productive dispatch (SQ_INSTS_VALU, active cells
# per launch of that run): profiles/stage_kernel_plain_cap_2048_counters.md.
# Issue peak: one wave64 VALU instruction per SIMD every 4 cycles, 1024
# SIMDs at 2.4 GHz.
PMC_VALU = {
    ("plain_cap_2048", "stage_kernel<2>"): (3.365e7, 1634048.0),
    ("plain_cap_2048", "stage_kernel<3>"): (3.286e7, 1634048.0),
    ("plain_cap_2048", "stage_kernel<4>"): (3.201e7, 1634048.0),
}
VALU_PEAK = 1024 * 2.4e9 / 4   # wave-instructions/s


def cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor()


def cpu_port(wl, budget_s: float = 20.0) -> dict:
    """The oracle (C restatement, 1 thread) on a bounded sample of the
    workload: its first steps from the same initial state, until budget_s."""
    from oracle.oracle import Oracle
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
        if el >= budget_s or steps >= wl.steps:
            break
    o.close()
    return {"value": cells / el, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"{wl.name}: first {steps} of {wl.steps} steps ({cells} cell-updates, "
                      f"{el:.1f} s) on {cpu_model()}, 1 thread"}


DECKS = os.path.join(ROOT, "tests", "golden", "workloads")


def _sample_deck(name: str, tend: float, work: str) -> str:
    """The workload's deck (tests/golden/workloads) ending at tend, one output."""
    import re
    with open(os.path.join(DECKS, name + ".txt")) as f:
        text = f.read()
    text = re.sub(r"(?m)^T end = .*$", f"T end = {tend!r}", text)
    text = re.sub(r"(?m)^N out = .*$", "N out = 1", text)
    path = os.path.join(work, name + "_sample.txt")
    with open(path, "w") as f:
        f.write(text)
    return path


def cpu_reference(wl, tend: float, copies: int = 1) -> dict | None:
    """The reference program (oracle/_ref/time/ref_run: its own Fortran
    sources built with release optimisation, serial) integrating the
    workload's deck to tend from its initial state, `copies` independent
    processes at once (one per core).  The cell-updates of the sample are
    those the device counts for the same integration (the same steps and
    active tiles).  Wall time includes the program's set-up and its two
    solution files.  None where the binary was not built."""
    import tempfile
    from kestrel_amd.runtime import Stepper
    from kestrel_amd import scenario as S
    from oracle import refrun
    if not refrun.available(refrun.TIME_BIN):
        return None
    g = Stepper(wl.cfg)
    g.set_topography(wl.b0v)
    g.set_state(wl.init)
    st = g.integrate_to(tend)
    g.close()
    with tempfile.TemporaryDirectory() as work:
        path = _sample_deck(wl.name, tend, work)
        b0 = None
        if wl.name.startswith("morpho_hills"):
            xv, yv = S.vertex_coords(wl.cfg)
            b0 = os.path.join(work, "b0.bin")
            refrun.write_b0(b0, wl.b0v, float(xv[0]), float(yv[0]), wl.cfg.dx, wl.cfg.dy)
        procs = []
        t0 = time.perf_counter()
        for k in range(copies):
            cmd = [refrun.TIME_BIN, path, os.path.join(work, f"out{k}"), "none"] + ([b0] if b0 else [])
            procs.append(subprocess.Popen(cmd, cwd=work, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL))
        rcs = [p.wait() for p in procs]
        el = time.perf_counter() - t0
    if any(rcs):
        return {"error": f"reference exited {rcs}"}
    cells = st["cell_updates"] * copies
    return {"value": cells / el, "unit": "cell-updates/s", "cores": copies, "kind": "reference",
            "sample": f"{wl.name} deck to t = {tend} s ({st['steps']} steps, {st['cell_updates']} "
                      f"cell-updates{' per process' if copies > 1 else ''}), "
                      f"{'%d concurrent processes' % copies if copies > 1 else '1 process'} of the "
                      f"reference built from its sources (-O2), {el:.1f} s wall on {cpu_model()}"}


def run(wl, steps: int, warmup: int):
    """Warm-up on its own handle, then K graph-launched steps from the initial
    state (value), then the same K steps again on a fresh handle with every
    kernel launch bracketed by HIP events (per-kernel launch times)."""
    import torch
    from kestrel_amd.runtime import Stepper

    def fresh():
        g = Stepper(wl.cfg)
        g.autotune(True)       # choices cached per configuration: the warm-up pass makes them
        g.set_topography(wl.b0v)
        g.set_state(wl.init)
        return g

    if warmup:
        w = fresh()
        w.step_n(warmup)
        w.close()
    g = fresh()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = g.step_n(steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    st["autotune"] = g.autotune_choices()
    g.close()
    g = fresh()
    g.profile(True)
    sp = g.step_n(steps)
    torch.cuda.synchronize()
    prof = g.profile_read()
    g.close()
    assert sp["cell_updates"] == st["cell_updates"], "profiled pass diverged from the timed pass"
    return st, el, prof


def algorithmic_bytes(name: str, mode: str, cells: float, redist_per_launch: float) -> float:
    """Algorithmic bytes of one launch of kernel `name` (FP64 words x 8)."""
    if name.startswith("stage_kernel<"):
        return STAGE_WORDS[mode][int(name[13]) - 1] * 8 * cells
    if name.startswith("morpho_kernel<"):
        return MORPHO_STAGE_WORDS * 8 * cells
    if name == "redistribute_kernel":
        return REDIST_WORDS * 8 * redist_per_launch
    return 0.0


def kernel_table(prof: dict, mode: str, cells: float, st: dict) -> list:
    nred = prof.get("redistribute_kernel", (0.0, 0))[1]
    rpl = st["redist_cells"] / nred if nred else 0.0
    total = sum(ms * n for ms, n in prof.values())
    rows = []
    for name, (ms, n) in sorted(prof.items(), key=lambda kv: -kv[1][0] * kv[1][1]):
        rows.append({"kernel": name, "launches": n, "mean_us": ms * 1e3,
                     "share": ms * n / total if total else 0.0,
                     "algorithmic_bytes": algorithmic_bytes(name, mode, cells, rpl)})
    return rows


def roofline(wl_name: str, rows: list, cells: float) -> dict:
    """HBM roofline of the dominant kernel (the most total time of the launches
    that did work) of the profiled pass."""
    top = rows[0]
    s = top["mean_us"] * 1e-6
    achieved = top["algorithmic_bytes"] / s / 1e9 if s > 0 else 0.0
    traffic = None
    base = top["kernel"].replace(" (H2)", "")
    key = (wl_name, base)
    if key in PMC_TRAFFIC:
        f, w, c0 = PMC_TRAFFIC[key]
        traffic = (f * PMC_FETCH_SCALE + w * PMC_WRITE_SCALE) * 1024 * cells / c0
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
           "kernel": top["kernel"], "algorithmic_bytes_per_launch": top["algorithmic_bytes"],
           "mean_launch_us": top["mean_us"], "cells_per_launch": cells}
    if key in PMC_VALU and s > 0:
        n, c0 = PMC_VALU[key]
        rate = n * cells / c0 / s
        out["valu_issue"] = {"achieved": rate, "peak": VALU_PEAK, "unit": "wave-instructions/s",
                             "frac": rate / VALU_PEAK}
    return out


def ensemble_leg(n: int = 64, steps: int = 100) -> dict:
    """The batched ensemble: n scenarios of `steps` steps each, advanced
    (a) one after another on the plain FP64 path, (b) together in one FP64
    batch, (c) together in one mixed-precision batch.  Setup (create,
    topography, state) is outside every timed region."""
    import torch
    from kestrel_amd import config as C
    from kestrel_amd import workloads
    from kestrel_amd.runtime import Batch, Stepper

    wls = workloads.ensemble(n, steps)

    def make(w, prec):
        w.cfg.precision = prec
        g = Stepper(w.cfg)
        g.set_topography(w.b0v)
        g.set_state(w.init)
        return g

    cells = 0
    el = 0.0
    for w in wls:
        g = make(w, C.PREC_FP64)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = g.step_n(steps)
        el += time.perf_counter() - t0
        cells += st["cell_updates"]
        g.close()
    out = {"workload": f"ensemble of {n}: sides log-uniform 128-1024, dx 5 m, cap U[1e3,1e5] m^3, "
                       f"Chezy U[0.005,0.05]", "steps_per_scenario": steps,
           "grid_cells": sum(w.cfg.NX * w.cfg.NY for w in wls), "cell_updates": cells,
           "plain_fp64_sequential": cells / el}
    for prec, name in ((C.PREC_FP64, "batched_fp64"), (C.PREC_MIXED, "batched_mixed")):
        b = Batch(n)
        gs = [make(w, prec) for w in wls]
        slots = [b.attach(g) for g in gs]
        for sl in slots:
            b.start(sl, nsteps=steps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        status = b.run()
        el = time.perf_counter() - t0
        c = sum(b.stats(sl)["cell_updates"] for sl in slots)
        assert c == cells and all(status[sl] == 1 for sl in slots), f"{name}: batch diverged"
        out[name] = c / el
        b.close()
        for g in gs:
            g.close()
    out["unit"] = "cell-updates/s"
    out["batched_mixed_vs_plain"] = out["batched_mixed"] / out["plain_fp64_sequential"]
    return out


def serving_leg(n: int = 32, rate: float = 4.0) -> dict:
    """A Poisson trace of n requests at `rate` per second (sides log-uniform
    64-2048, T ~ U[60, 600] s simulated) served with continuous batching
    under a 200 GB HBM budget, mixed precision."""
    from kestrel_amd import serving
    reqs = serving.poisson_trace(n, rate)
    srv = serving.Server(256, 200e9)
    out = srv.run(reqs)
    srv.close()
    out["trace"] = f"{n} requests, Poisson rate {rate}/s, sides 64-2048, T U[60,600] s, mixed precision"
    return out


def volume_leg(calls: int = 200) -> dict:
    """kst_volume (CalculateVolume on the device) on the plain 2048x2048 cap
    after 10 steps: wall time per call, which includes the control-block
    read-back, both reduction kernels and the copy of the result; the
    volume kernel's own duration is in the committed rocprofv3 summary."""
    from kestrel_amd import workloads
    from kestrel_amd.runtime import Stepper
    wl = workloads.plain_cap_2048(10)
    g = Stepper(wl.cfg)
    g.set_topography(wl.b0v)
    g.set_state(wl.init)
    st = g.step_n(10)
    for _ in range(5):
        g.volume()
    t0 = time.perf_counter()
    for _ in range(calls):
        v = g.volume()
    el = (time.perf_counter() - t0) / calls
    g.close()
    cells = st["active_tiles"] * wl.cfg.nx_per_tile * wl.cfg.ny_per_tile
    return {"workload": wl.name, "active_cells": cells, "bytes_per_cell": 40,
            "call_ms": el * 1e3, "volume_m3": float(v[0])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="morpho_hills_1024")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--volume-only", action="store_true",
                    help="only the kst_volume leg (for profiling the volume kernels)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is designed, not built)"}))
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)
    if args.volume_only:
        print(json.dumps({"volume": volume_leg()}), flush=True)
        return
    from kestrel_amd import workloads

    wl = getattr(workloads, args.workload)()
    wl.steps = args.steps
    st, el, prof = run(wl, args.steps, args.warmup)
    cells = st["cell_updates"]
    value = cells / el
    cells_per_step = cells / max(1, st["steps"])
    mode = "morpho" if wl.cfg.morphodynamics else "hydraulic"
    rows = kernel_table(prof, mode, cells_per_step, st)
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
                   "operator": "Strang H(dt) M(2dt) H(dt)" if mode == "morpho" else "hydraulic",
                   "active_cells_per_step": cells_per_step, "final_active_tiles": st["active_tiles"],
                   "simulated_time_s": st["t"], "precision": "fp64 plain",
                   "rhs_evals": st["rhs_evals"], "rejected": st["rejected"],
                   "redistributed_cells": st["redist_cells"], "autotune": st["autotune"]},
        "roofline": roofline(wl.name, rows, cells_per_step),
        "kernels": [{k: (round(v, 3) if isinstance(v, float) else v) for k, v in r.items()} for r in rows],
        "step_roofline_frac": value * STEP_WORDS[mode] * 8 / (HBM_PEAK_GBS * 1e9),
    }
    if not args.no_plain and args.workload != "plain_cap_2048":
        pw = workloads.plain_cap_2048(200)
        pst, pel, pprof = run(pw, 200, 10)
        pcells = pst["cell_updates"] / max(1, pst["steps"])
        prows = kernel_table(pprof, "hydraulic", pcells, pst)
        res["plain_fp64"] = {"workload": pw.name, "grid": "2048x2048", "steps": 200,
                             "value": pst["cell_updates"] / pel, "unit": "cell-updates/s",
                             "ms_per_step": pel * 1e3 / 200, "roofline": roofline(pw.name, prows, pcells),
                             "step_roofline_frac": pst["cell_updates"] / pel * STEP_WORDS["hydraulic"] * 8
                             / (HBM_PEAK_GBS * 1e9), "autotune": pst["autotune"]}
        # the same deck in mixed precision (FP32 RHS, stage increments and
        # non-hydrostatic face arithmetic; FP64 u^n and free surface)
        from kestrel_amd import config as C
        mw = workloads.plain_cap_2048(200)
        mw.cfg.precision = C.PREC_MIXED
        mst, mel, _ = run(mw, 200, 10)
        res["plain_fp64"]["mixed"] = {"value": mst["cell_updates"] / mel, "unit": "cell-updates/s",
                                      "ms_per_step": mel * 1e3 / 200,
                                      "vs_fp64": (mst["cell_updates"] / mel) / res["plain_fp64"]["value"]}
        if not args.no_cpu:
            # two steps of the same deck in the reference (about 10 s of CPU work)
            res["plain_fp64"]["cpu_baseline"] = cpu_reference(pw, 0.0225)
    if not args.no_plain:
        res["volume"] = volume_leg()
    if not args.no_batch:
        res["ensemble"] = ensemble_leg()
        if "plain_fp64" in res:
            # the batched ensemble against the plain FP64 path on the full chip
            res["ensemble"]["batched_mixed_vs_plain_2048"] = (res["ensemble"]["batched_mixed"]
                                                             / res["plain_fp64"]["value"])
    if not args.no_serve:
        res["serving"] = serving_leg(args.serve_requests)
    if not args.no_cpu:
        ref = None
        if wl.name == "morpho_hills_1024":
            ref = cpu_reference(wl, 8.0)
        elif wl.name in ("plain_cap_256", "plain_cap_2048"):
            ref = cpu_reference(wl, 2.8125 if wl.name == "plain_cap_256" else 0.0225)
        port = cpu_port(wl, args.cpu_budget)
        res["cpu_baseline"] = ref if ref and "error" not in ref else port
        res["cpu_port"] = port
        # all host cores: independent reference processes of the 256 x 256 cap
        # release (the reference is serial), one per core up to 16
        from kestrel_amd import workloads as W
        ncore = min(16, os.cpu_count() or 1)
        res["cpu_all_cores"] = cpu_reference(W.plain_cap(256, 250), 2.8125, copies=ncore)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
