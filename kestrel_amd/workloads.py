"""The measured workloads, with their sizes (the synthetic decks of the
reference's time step; every value here is a Kestrel input-deck setting).

plain_cap_256      : hydraulic cap release, 256x256 cells, dx = 1 m, plane
                     inclined 10 degrees, paraboloid cap (1 m, radius 20 m),
                     Chezy 0.01, CFL 0.25, max dt 0.05 s (so dt = 0.01125 s
                     while inactive tiles remain), 1000 steps, FP64.
plain_cap(n)       : the same physics on an n x n grid (n = 2048 for the
                     precision study), cap scaled to stay clear of the edge.
morpho_hills_1024  : 1024x1024, dx = 5 m, seeded 32 Gaussian hills on a
                     5 degree ramp, one-cell flux source q = 20 m^2/s (Q = q*dx),
                     psi = 0.3, Pouliquen drag, Mixed erosion, Spearman-Manning
                     deposition, eddy viscosity 1e-2, 2000 steps.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import config as C
from . import scenario as S


@dataclass
class Workload:
    name: str
    cfg: C.KstConfig
    b0v: np.ndarray
    init: S.InitialState
    steps: int
    caps: list
    sources: list


def plain_cap(n: int = 256, steps: int = 1000, tile: int = 32, dx: float = 1.0,
              cap_x: float | None = None, cap_radius: float | None = None) -> Workload:
    """Plain hydraulic 2D cap release on a 10 degree x-plane."""
    nt = n // tile
    cfg = C.default_config(nt, nt, tile, tile, dx=dx)
    C.hydraulic_only(cfg)
    cfg.drag = C.DRAGS["chezy"]
    cfg.chezy_co = 0.01
    cfg.cfl = 0.25
    cfg.max_dt = 0.05
    C.derive(cfg)
    b0v = S.topo_xslope(cfg, math.tan(math.radians(10.0)))
    L = n * dx
    r = 20.0 if cap_radius is None else cap_radius
    x0 = (0.25 * L) if cap_x is None else cap_x
    caps = [S.Cap(x0, 0.0, r, 1.0, shape="para")]
    init = S.initial_state(cfg, b0v, caps)
    return Workload(f"plain_cap_{n}", cfg, b0v, init, steps, caps, [])


def plain_cap_2048(steps: int = 200) -> Workload:
    """The plain hydraulic physics (10 degree plane, Chezy 0.01, CFL 0.25,
    max dt 0.05 s) at 2048x2048 cells, dx = 1 m, with a paraboloid cap of
    height 1 m and radius 700 m centred at x = +200 m: about 1.6 M cells in
    about 1550 active tiles, so the step is measured on a full chip."""
    return plain_cap(2048, steps, tile=32, dx=1.0, cap_x=200.0, cap_radius=700.0)


def plain_cap_256_sample() -> Workload:
    return plain_cap(256, 1000)


def fixture_case(name: str) -> Workload:
    """Small 40x40 decks (5x5 tiles of 8 cells, dx = 1 m) exercising each
    runtime-selected closure, for the reference-vector fixtures."""
    cfg = C.default_config(5, 5, 8, 8, dx=1.0)
    C.hydraulic_only(cfg)
    cfg.cfl = 0.25
    cfg.max_dt = 0.05
    slope = math.tan(math.radians(10.0))
    psi = 0.0
    if name == "minmod1":
        cfg.limiter = C.LIMITERS["minmod1"]
    elif name == "vanalbada":
        cfg.limiter = C.LIMITERS["vanalbada"]
    elif name == "weno":
        cfg.limiter = C.LIMITERS["weno"]
    elif name == "nolimiter":
        cfg.limiter = C.LIMITERS["none"]
    elif name in ("coulomb", "voellmy", "pouliquen", "edwards2019", "manning", "variable"):
        cfg.drag = C.DRAGS[name]
        psi = 0.3 if name == "variable" else 0.0
    elif name == "flat_nogeo":
        cfg.geometric_factors = 0
        slope = 0.0
    elif name == "eddy":
        cfg.eddy_viscosity = 0.05
    elif name == "psi":
        psi = 0.4
    elif name != "chezy":
        raise ValueError(name)
    C.derive(cfg)
    b0v = S.topo_xslope(cfg, slope) if slope else S.topo_flat(cfg)
    caps = [S.Cap(0.5, -0.5, 6.0, 1.0, psi=psi, shape="para")]
    init = S.initial_state(cfg, b0v, caps)
    return Workload(f"fixture_{name}", cfg, b0v, init, 0, caps, [])


FIXTURE_CASES = ["chezy", "minmod1", "vanalbada", "weno", "nolimiter", "coulomb", "voellmy",
                 "pouliquen", "edwards2019", "manning", "variable", "flat_nogeo", "eddy", "psi"]
