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
ensemble           : 64 hydraulic scenarios, sides log-uniform in [128, 1024]
                     (multiples of 32), dx = 5 m, seeded Gaussian hills on a
                     2 degree ramp, a paraboloid cap of volume U[1e3, 1e5] m^3
                     (radius 50 m), Chezy coefficient U[0.005, 0.05].
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
    C.reference_defaults(cfg, {"chezy_co"})
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


def morpho_hills(n: int = 1024, steps: int = 2000, tile: int = 32, dx: float = 5.0,
                 seed: int = 20240611, q: float = 20.0, psi: float = 0.3,
                 src_x: float | None = None, src_y: float = 0.0, n_hills: int = 32,
                 drag: str = "pouliquen") -> Workload:
    """Full morphodynamics: an n x n grid of dx cells over a seeded sum of
    n_hills Gaussian hills (amplitude U[5, 50] m, width U[50, 300] m) on a
    5 degree x-ramp, fed by a one-cell flux source (radius 0.4 dx at a cell
    centre) of Q = q*dx m^3/s with solids fraction psi, Pouliquen drag, Mixed
    erosion, Spearman-Manning deposition, smooth erosion transition, tanh
    damping, eddy viscosity 1e-2 m^2/s.  The flow runs toward -x."""
    nt = n // tile
    cfg = C.default_config(nt, nt, tile, tile, dx=dx)
    cfg.drag = C.DRAGS[drag]
    cfg.erosion = C.EROSIONS["mixed"]
    cfg.deposition = C.DEPOSITIONS["spearman manning"]
    cfg.morphodynamics = 1
    cfg.eddy_viscosity = 1e-2
    # as the deck gives them: Chezy Co and the Voellmy switch explicitly
    C.reference_defaults(cfg, {"chezy_co", "voellmy_switch_rate", "voellmy_switch_value"})
    C.derive(cfg)
    hills = S.Hills(seed, n=n_hills)
    b0v = S.topo_gaussian_hills(cfg, hills)
    L = n * dx
    x0 = 0.25 * L if src_x is None else src_x
    # snap to a cell centre
    xc, yc = S.cell_centres(cfg)
    x0 = float(xc[np.argmin(np.abs(xc - x0))])
    y0 = float(yc[np.argmin(np.abs(yc - src_y))])
    src = C.FluxSource(x0, y0, 0.4 * dx, [0.0, 1.0e9], [q * dx, q * dx], [psi, psi])
    init = S.initial_state(cfg, b0v, [], [src])
    C.set_sources(cfg, [src])
    return Workload(f"morpho_hills_{n}", cfg, b0v, init, steps, [], [src])


def morpho_hills_1024(steps: int = 2000) -> Workload:
    """The full morphodynamic run: 1024x1024 cells of 5 m (32x32 tiles of 32),
    2000 Strang steps."""
    return morpho_hills(1024, steps)


def morpho_cap(n: int = 128, steps: int = 100, tile: int = 16, dx: float = 1.0,
               psi: float = 0.3, radius: float = 12.0, height: float = 1.0,
               slope_deg: float = 10.0, u: float = 0.0, drag: str = "chezy",
               erosion: str = "mixed", deposition: str = "spearman manning",
               transition: str = "smooth", damping: str = "tanh", ero_rate: float | None = None,
               settling_speed: float | None = None, ero_critical_height: float | None = None) -> Workload:
    """Small morphodynamic decks for parity tests: a paraboloid cap with solids
    fraction psi released on an x-plane (centred at x = 0), with selectable
    closures."""
    nt = n // tile
    cfg = C.default_config(nt, nt, tile, tile, dx=dx)
    cfg.drag = C.DRAGS[drag]
    cfg.erosion = C.EROSIONS[erosion]
    cfg.deposition = C.DEPOSITIONS[deposition]
    cfg.ero_transition = C.TRANSITIONS[transition]
    cfg.morpho_damping = C.DAMPINGS[damping]
    cfg.morphodynamics = 1
    cfg.cfl = 0.25
    cfg.max_dt = 0.05
    if ero_rate is not None:
        cfg.ero_rate = ero_rate
    if ero_critical_height is not None:
        cfg.ero_critical_height = ero_critical_height
    C.derive(cfg, settling_speed=settling_speed)
    b0v = S.topo_xslope(cfg, math.tan(math.radians(slope_deg))) if slope_deg else S.topo_flat(cfg)
    caps = [S.Cap(0.0, 0.0, radius, height, psi=psi, u=u, shape="para")]
    init = S.initial_state(cfg, b0v, caps)
    return Workload(f"morpho_cap_{n}", cfg, b0v, init, steps, caps, [])


FIXTURE_CASES = ["chezy", "minmod1", "vanalbada", "weno", "nolimiter", "coulomb", "voellmy",
                 "pouliquen", "edwards2019", "manning", "variable", "flat_nogeo", "eddy", "psi"]


def morpho_fixture_case(name: str) -> Workload:
    """Small 40x40 morphodynamic decks (5x5 tiles of 8 cells, dx = 1 m): a
    paraboloid cap (radius 6 m, height 1 m, psi 0.3) released on a 10 degree
    x-plane, one deck per erosion/deposition/transition/damping choice."""
    kw = dict(n=40, tile=8, radius=6.0, psi=0.3)
    if name == "mixed":
        pass
    elif name == "simple":
        kw.update(erosion="simple", deposition="simple", transition="step", damping="rat3")
    elif name == "fluid":
        kw.update(erosion="fluid", deposition="none", transition="none", damping="none")
    elif name == "granular":
        kw.update(erosion="granular", transition="step", damping="tanh")
    elif name == "fastsettling":
        kw.update(settling_speed=0.2, ero_critical_height=0.05)
    elif name == "nogeo":
        kw.update(slope_deg=0.0, u=1.0)
    else:
        raise ValueError(name)
    wl = morpho_cap(**kw)
    if name == "nogeo":
        wl.cfg.geometric_factors = 0
    wl.name = f"morpho_fixture_{name}"
    return wl


MORPHO_FIXTURE_CASES = ["mixed", "simple", "fluid", "granular", "fastsettling", "nogeo"]


def ensemble_member(k: int, seed: int = 20240612, side_range=(128, 1024), dx: float = 5.0,
                    tile: int = 16, steps: int = 100, precision: int = C.PREC_MIXED,
                    n_hills: int = 16) -> Workload:
    """Scenario k of the batched ensemble: its own seeded draw of grid side
    (log-uniform in side_range, a multiple of 32), Gaussian-hill DEM on a
    2 degree x-ramp, cap volume V ~ U[1e3, 1e5] m^3 painted at the domain
    centre as a paraboloid of radius 0.06 L (height 2 V / (pi r^2), as
    capVolume sets it, src/InitConds.f90:362-368), and Chezy coefficient
    U[0.005, 0.05].  Tiles of 16 cells."""
    rng = np.random.default_rng([seed, k])
    lo, hi = math.log(side_range[0]), math.log(side_range[1])
    side = int(round(math.exp(rng.uniform(lo, hi)) / 32)) * 32
    side = min(max(side, side_range[0]), side_range[1])
    vol = float(rng.uniform(1.0e3, 1.0e5))
    chezy = float(rng.uniform(0.005, 0.05))
    hseed = int(rng.integers(1 << 30))
    nt = side // tile
    cfg = C.default_config(nt, nt, tile, tile, dx=dx)
    C.hydraulic_only(cfg)
    cfg.drag = C.DRAGS["chezy"]
    cfg.chezy_co = chezy
    cfg.cfl = 0.25
    cfg.precision = precision
    C.derive(cfg)
    b0v = S.topo_gaussian_hills(cfg, S.Hills(hseed, n=n_hills, ramp_deg=2.0))
    L = side * dx
    cap_radius = 0.06 * L
    height = 2.0 * vol / (math.pi * cap_radius * cap_radius)
    caps = [S.Cap(0.0, 0.0, cap_radius, height, shape="para")]
    init = S.initial_state(cfg, b0v, caps)
    return Workload(f"ensemble_{k}_side{side}", cfg, b0v, init, steps, caps, [])


def ensemble(n: int = 64, steps: int = 100, precision: int = C.PREC_MIXED, **kw) -> list:
    """The batched ensemble: n independent scenarios (see ensemble_member)."""
    return [ensemble_member(k, steps=steps, precision=precision, **kw) for k in range(n)]
