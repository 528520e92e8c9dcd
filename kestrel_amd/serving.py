"""Serving: hazard scenarios arrive over time and run together on one GPU.

A Server owns one Batch (a fixed table of slots advanced by one launch per
kernel, kst_batch_*).  Requests wait in arrival order; between batch runs of
`poll_iters` iterations the server retires the scenarios that finished
(their state stays on the device until then), frees their memory, and admits
waiting requests while a slot is free and their device memory fits in the
HBM budget.  Work of every resident scenario overlaps in the same launches:
a scenario that arrives while others run joins the next launch instead of
waiting for them (continuous batching).  Each scenario's result is the same
bits as running it alone.

    python -m kestrel_amd.serving [--requests N] [--rate 4] [--budget-gb 200]
"""
from __future__ import annotations

import argparse
import heapq
import json
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import config as C
from . import scenario as S
from . import workloads
from .runtime import SLOT_DONE, SLOT_ERROR, SLOT_RUNNING, Batch, KestrelError, Stepper

# Device bytes per cell of a scenario (every tile of the grid), by precision:
# the FP64 state, stages and RHS, the cell and face topography, ghost slopes,
# maxima and tile lists (kst_create's allocations).
BYTES_PER_CELL = {C.PREC_FP64: 8 * 48, C.PREC_MIXED: 8 * 45}


def device_bytes(cfg) -> int:
    return int(cfg.NX * cfg.NY * BYTES_PER_CELL[cfg.precision])


@dataclass
class Request:
    rid: int
    arrival: float                 # seconds after the trace starts
    workload: workloads.Workload
    tend: float                    # simulated time to reach
    # filled in by the server
    admitted: float = math.nan
    finished: float = math.nan
    status: str = "queued"
    stats: dict = field(default_factory=dict)
    state: dict | None = None


def trace_request(k: int, seed: int = 20240613, side_range=(64, 2048),
                  precision: int = C.PREC_MIXED, t_range=(60.0, 600.0)) -> tuple:
    """Request k of the serving trace: a grid side log-uniform in side_range
    (a multiple of 32) with dx = 5 m, coarsened by powers of two below 512
    cells so that every domain spans at least 2.5 km; a seeded Gaussian-hill
    DEM in a shallow bowl (50 m rise to the domain edge) that keeps the flow
    on the grid; a paraboloid cap of volume V ~ U[1e3, 1e5] m^3 (radius
    0.06 L) or a one-cell flux source releasing the same kind of volume over
    the whole run (Q = V / T); Chezy coefficient U[0.005, 0.05]; run to
    T ~ U[t_range] s."""
    rng = np.random.default_rng([seed, k])
    lo, hi = math.log(side_range[0]), math.log(side_range[1])
    side = int(round(math.exp(rng.uniform(lo, hi)) / 32)) * 32
    side = min(max(side, max(64, side_range[0])), side_range[1])
    tend = float(rng.uniform(*t_range))
    chezy = float(rng.uniform(0.005, 0.05))
    use_flux = bool(rng.uniform() < 0.5)
    vol = float(rng.uniform(1.0e3, 1.0e5))
    hseed = int(rng.integers(1 << 30))
    dx = 5.0 * 2 ** max(0, math.ceil(math.log2(512 / side)))
    tile = 16
    nt = side // tile
    cfg = C.default_config(nt, nt, tile, tile, dx=dx)
    C.hydraulic_only(cfg)
    cfg.drag = C.DRAGS["chezy"]
    cfg.chezy_co = chezy
    cfg.cfl = 0.25
    cfg.precision = precision
    C.derive(cfg)
    L = side * dx
    b0v = S.topo_gaussian_hills(cfg, S.Hills(hseed, n=16, ramp_deg=0.0))
    xv, yv = S.vertex_coords(cfg)
    X, Y = np.meshgrid(xv, yv)
    b0v = b0v + 50.0 * (X * X + Y * Y) / (0.5 * L) ** 2
    if use_flux:
        xc, yc = S.cell_centres(cfg)
        x0 = float(xc[np.argmin(np.abs(xc))])
        y0 = float(yc[np.argmin(np.abs(yc))])
        Q = vol / tend
        src = C.FluxSource(x0, y0, 0.4 * dx, [0.0, 1.0e9], [Q, Q], [0.0, 0.0])
        init = S.initial_state(cfg, b0v, [], [src])
        C.set_sources(cfg, [src])
        wl = workloads.Workload(f"serve_{k}_side{side}_flux", cfg, b0v, init, 0, [], [src])
    else:
        r = 0.06 * L
        caps = [S.Cap(0.0, 0.0, r, 2.0 * vol / (math.pi * r * r), shape="para")]
        init = S.initial_state(cfg, b0v, caps)
        wl = workloads.Workload(f"serve_{k}_side{side}_cap", cfg, b0v, init, 0, caps, [])
    return wl, tend


def poisson_trace(n: int, rate: float = 4.0, seed: int = 20240613, **kw) -> list:
    """n requests with Poisson arrivals at `rate` per second."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for k in range(n):
        t += float(rng.exponential(1.0 / rate))
        wl, tend = trace_request(k, seed=seed, **kw)
        reqs.append(Request(k, t, wl, tend))
    return reqs


class Server:
    def __init__(self, max_slots: int = 256, budget_bytes: float = 200e9, poll_iters: int = 32,
                 keep_state: bool = False):
        self.batch = Batch(max_slots)
        self.max_slots = max_slots
        self.budget = budget_bytes
        self.poll_iters = poll_iters
        self.keep_state = keep_state
        self.used = 0
        self.queue: list = []          # (arrival, rid, Request)
        self.resident: dict = {}       # slot -> (Request, Stepper, bytes)
        self.done: list = []
        self.cell_updates = 0
        self.peak_resident = 0
        self.peak_bytes = 0

    def submit(self, req: Request):
        heapq.heappush(self.queue, (req.arrival, req.rid, req))

    def _admit(self, now: float):
        while self.queue and self.queue[0][0] <= now and len(self.resident) < self.max_slots:
            req = self.queue[0][2]
            need = device_bytes(req.workload.cfg)
            if self.used + need > self.budget and self.resident:
                break                      # wait for memory
            heapq.heappop(self.queue)
            g = Stepper(req.workload.cfg)
            g.set_topography(req.workload.b0v)
            g.set_state(req.workload.init)
            slot = self.batch.attach(g)
            self.batch.start(slot, tend=req.tend)
            req.admitted = now
            req.status = "running"
            self.resident[slot] = (req, g, need)
            self.used += need
            self.peak_resident = max(self.peak_resident, len(self.resident))
            self.peak_bytes = max(self.peak_bytes, self.used)

    def _retire(self, status: list, now: float):
        for slot in list(self.resident):
            if status[slot] == SLOT_RUNNING:
                continue
            req, g, need = self.resident.pop(slot)
            try:
                req.stats = self.batch.stats(slot)
                req.status = "done" if status[slot] == SLOT_DONE else "error"
            except KestrelError as e:
                req.status = f"error: {e}"
            self.cell_updates += req.stats.get("cell_updates", 0)
            if self.keep_state:
                req.state = g.state()
            req.finished = now
            self.batch.detach(slot)
            g.close()
            self.used -= need
            self.done.append(req)

    def run(self, requests: list, clock=time.perf_counter) -> dict:
        """Serve the requests (arrival times relative to the call) to the end."""
        for r in requests:
            self.submit(r)
        t0 = clock()
        while self.queue or self.resident:
            now = clock() - t0
            self._admit(now)
            if not self.resident:
                # idle until the next arrival
                wait = self.queue[0][0] - (clock() - t0)
                if wait > 0:
                    time.sleep(wait)
                continue
            status = self.batch.run(self.poll_iters)
            self._retire(status, clock() - t0)
        wall = clock() - t0
        lat = np.array([r.finished - r.arrival for r in self.done])
        ok = [r for r in self.done if r.status == "done"]
        return {"requests": len(self.done), "completed": len(ok), "wall_s": wall,
                "peak_resident": self.peak_resident, "peak_device_bytes": self.peak_bytes,
                "cell_updates": self.cell_updates,
                "cell_updates_per_s": self.cell_updates / wall if wall > 0 else 0.0,
                "latency_p50_s": float(np.percentile(lat, 50)) if lat.size else math.nan,
                "latency_p99_s": float(np.percentile(lat, 99)) if lat.size else math.nan}

    def close(self):
        for slot, (req, g, need) in list(self.resident.items()):
            self.batch.detach(slot)
            g.close()
        self.resident.clear()
        self.batch.close()


def main(argv=None):
    ap = argparse.ArgumentParser(description="Serve a Poisson trace of hazard scenarios on one GPU.")
    ap.add_argument("--requests", type=int, default=32)
    ap.add_argument("--rate", type=float, default=4.0)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--slots", type=int, default=256)
    ap.add_argument("--max-side", type=int, default=2048)
    ap.add_argument("--t-max", type=float, default=600.0)
    a = ap.parse_args(argv)
    reqs = poisson_trace(a.requests, a.rate, side_range=(64, a.max_side), t_range=(60.0, a.t_max))
    srv = Server(a.slots, a.budget_gb * 1e9)
    out = srv.run(reqs)
    srv.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
