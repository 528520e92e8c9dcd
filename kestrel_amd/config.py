"""Host-side mirror of Kestrel's run settings for the time-step boundary.

``KstConfig`` is the ctypes image of ``kst_config`` (include/kestrel_step.h).
``default_config`` fills the reference defaults (src/Parameters.f90:41-77,
src/SolverSettings.f90:42-51) and ``derive`` the derived parameters of
src/Parameters.f90:629-648.  Closure and limiter names follow the input-deck
spellings bound in src/Parameters.f90:175-420 and src/SolverSettings.f90:102-126.
"""
from __future__ import annotations

import ctypes
import math
from typing import Iterable

BC_HALT, BC_PERIODIC, BC_DIRICHLET, BC_SPONGE = 0, 1, 2, 3
BCS = {"halt": BC_HALT, "periodic": BC_PERIODIC, "dirichlet": BC_DIRICHLET, "sponge": BC_SPONGE}

LIMITERS = {"minmod1": 0, "minmod2": 1, "minmod": 1, "vanalbada": 2, "weno": 3, "none": 4}
DRAGS = {"chezy": 0, "coulomb": 1, "voellmy": 2, "pouliquen": 3, "edwards2019": 4,
         "manning": 5, "variable": 6}
EROSIONS = {"off": 0, "none": 0, "simple": 1, "fluid": 2, "granular": 3, "mixed": 4}
DEPOSITIONS = {"off": 0, "none": 0, "simple": 1, "spearman manning": 2}
TRANSITIONS = {"none": 0, "off": 0, "smooth": 1, "step": 2}
DAMPINGS = {"none": 0, "off": 0, "tanh": 1, "rat3": 2}
SWITCHES = {"tanh": 0, "rat3": 1, "cos": 2, "equal": 3, "zero": 4, "one": 5, "linear": 6, "step": 7}

PREC_FP64 = 0
PREC_MIXED = 1

TILE_OFF, TILE_ACTIVE, TILE_GHOST = 0, 1, 2

# Kinematic viscosity of water used by the derived settling parameters
# (src/Parameters.f90:76).
VISC_W = 1.2e-6


class KstFluxSource(ctypes.Structure):
    _fields_ = [
        ("x", ctypes.c_double), ("y", ctypes.c_double), ("radius", ctypes.c_double),
        ("n_series", ctypes.c_int32), ("num_cells", ctypes.c_int32),
        ("time", ctypes.POINTER(ctypes.c_double)),
        ("flux", ctypes.POINTER(ctypes.c_double)),
        ("psi", ctypes.POINTER(ctypes.c_double)),
    ]


_INT_FIELDS = ["nx_tiles", "ny_tiles", "nx_per_tile", "ny_per_tile"]
_DBL_GRID = ["dx", "dy", "x_size", "y_size"]
_INT_OPTS = ["bc", "limiter", "geometric_factors", "drag", "erosion", "deposition",
             "ero_transition", "morpho_damping", "fswitch", "morphodynamics", "tile_buffer",
             "track_maxima", "precision", "n_sources"]
_DBL_PARAMS = ["g", "rhow", "rhos", "chezy_co", "coulomb_co", "manning_co",
               "pouliquen_min", "pouliquen_max", "pouliquen_intermediate", "pouliquen_beta",
               "edwards_betastar", "edwards_kappa", "edwards_gamma",
               "ero_rate", "ero_rate_granular", "ero_depth", "ero_critical_height",
               "voellmy_switch_rate", "voellmy_switch_value",
               "max_pack", "bed_porosity", "solid_diameter", "eddy_viscosity",
               "cfl", "max_dt", "height_threshold",
               "gred", "rep", "ws0", "nsettling", "critical_shields", "diffusive_time_scale"]


class KstConfig(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_int32) for n in _INT_FIELDS]
                + [(n, ctypes.c_double) for n in _DBL_GRID]
                + [(n, ctypes.c_int32) for n in _INT_OPTS]
                + [(n, ctypes.c_double) for n in _DBL_PARAMS]
                + [("sources", ctypes.POINTER(KstFluxSource))]
                + [(n, ctypes.c_double) for n in ("bc_hn", "bc_u", "bc_v", "bc_psi", "sponge_strength")])

    # Keep the flux-source arrays alive as long as the struct.
    _keep: list

    @property
    def NX(self) -> int:
        return self.nx_tiles * self.nx_per_tile

    @property
    def NY(self) -> int:
        return self.ny_tiles * self.ny_per_tile

    @property
    def ntiles(self) -> int:
        return self.nx_tiles * self.ny_tiles

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "sources"}


class FluxSource:
    """One flux source: location, radius and (time, flux, psi) knots."""

    def __init__(self, x: float, y: float, radius: float, time: Iterable[float],
                 flux: Iterable[float], psi: Iterable[float]):
        self.x, self.y, self.radius = float(x), float(y), float(radius)
        self.time = [float(v) for v in time]
        self.flux = [float(v) for v in flux]
        self.psi = [float(v) for v in psi]
        if not (len(self.time) == len(self.flux) == len(self.psi) >= 1):
            raise ValueError("flux source series must have equal, non-zero lengths")
        self.num_cells = 0


def default_config(nx_tiles: int, ny_tiles: int, nx_per_tile: int = 32, ny_per_tile: int = 32,
                   dx: float = 1.0, dy: float | None = None) -> KstConfig:
    """A 2D configuration with the reference defaults."""
    c = KstConfig()
    c.nx_tiles, c.ny_tiles = nx_tiles, ny_tiles
    c.nx_per_tile, c.ny_per_tile = nx_per_tile, ny_per_tile
    c.dx = float(dx)
    c.dy = float(dx if dy is None else dy)
    c.x_size = c.dx * nx_tiles * nx_per_tile
    c.y_size = c.dy * ny_tiles * ny_per_tile
    c.bc = BC_HALT
    c.limiter = LIMITERS["minmod2"]
    c.geometric_factors = 1
    c.drag = DRAGS["chezy"]
    # The reference defaults to Mixed erosion with morphodynamics on when the
    # deck names no erosion law (src/Parameters.f90:576-580).
    c.erosion = EROSIONS["mixed"]
    c.deposition = DEPOSITIONS["spearman manning"]
    c.ero_transition = TRANSITIONS["smooth"]
    c.morpho_damping = DAMPINGS["tanh"]
    c.fswitch = SWITCHES["tanh"]
    c.morphodynamics = 1
    c.tile_buffer = 1
    c.track_maxima = 1
    c.precision = PREC_FP64
    c.n_sources = 0
    c.g = 9.81
    c.rhow, c.rhos = 1000.0, 2000.0
    c.chezy_co, c.coulomb_co, c.manning_co = 0.01, 0.1, 0.03
    c.pouliquen_min, c.pouliquen_max = 0.1, 0.4
    c.pouliquen_intermediate, c.pouliquen_beta = 0.2, 0.136
    c.edwards_betastar, c.edwards_kappa, c.edwards_gamma = 0.136, 1.0, 0.0
    c.ero_rate, c.ero_rate_granular, c.ero_depth = 0.001, 4.0, 1.0
    c.ero_critical_height = 0.01
    c.voellmy_switch_rate, c.voellmy_switch_value = 3.0, 0.2
    c.max_pack, c.bed_porosity, c.solid_diameter = 0.65, 0.35, 1e-3
    c.eddy_viscosity = 0.0
    c.cfl = 0.25
    c.max_dt = 1.7976931348623157e308
    c.height_threshold = 1e-6
    c.sponge_strength = 0.2
    c._keep = []
    derive(c)
    return c


def hydraulic_only(c: KstConfig) -> KstConfig:
    """Erosion = Off: morphodynamics off (the reference's 'Erosion = Off')."""
    c.erosion = EROSIONS["off"]
    c.morphodynamics = 0
    return c


# Drag and erosion parameters: the reference gives each a default only when
# the chosen closure is one that it names for it (src/Parameters.f90:494-596);
# the others keep the value of a fresh RunSet, zero.  Closures may still read
# them (FlowShieldsNumber reads ChezyCo, src/Closures.f90:221; the switch
# functions read the Voellmy switch rate and value, src/Closures.f90:795-915).
_DRAG_DEFAULTS = {
    "chezy_co": (("chezy", "voellmy"), 0.01),
    "coulomb_co": (("coulomb", "voellmy"), 0.1),
    "manning_co": (("manning",), 0.03),
    "pouliquen_min": (("pouliquen", "variable", "edwards2019"), 0.1),
    "pouliquen_max": (("pouliquen", "variable", "edwards2019"), 0.4),
    "pouliquen_beta": (("pouliquen", "variable", "edwards2019"), 0.136),
    "pouliquen_intermediate": (("edwards2019",), 0.2),
    "edwards_betastar": (("edwards2019",), 0.136),
    "edwards_kappa": (("edwards2019",), 1.0),
    "edwards_gamma": (("edwards2019",), 0.0),
    "voellmy_switch_rate": (("variable",), 3.0),
    "voellmy_switch_value": (("variable",), 0.2),
}
_EROSION_DEFAULTS = {
    "ero_rate": (("fluid", "mixed"), 0.001),
    "ero_rate_granular": (("granular", "mixed"), 4.0),
    "ero_depth": (("simple", "fluid", "granular", "mixed"), 1.0),
    "ero_critical_height": (("simple", "fluid", "granular", "mixed"), 0.01),
}


def reference_defaults(c: KstConfig, given=()) -> KstConfig:
    """Drag and erosion parameters as the reference leaves them after reading
    a deck that sets the fields in `given` (src/Parameters.f90:494-596): the
    chosen closures' defaults, zero for the rest."""
    drag = {v: k for k, v in DRAGS.items()}[c.drag]
    ero = {v: k for k, v in EROSIONS.items() if k not in ("none",)}[c.erosion]
    for table, choice in ((_DRAG_DEFAULTS, drag), (_EROSION_DEFAULTS, ero)):
        for name, (users, default) in table.items():
            if name in given:
                continue
            setattr(c, name, default if choice in users else 0.0)
    return c


def derive(c: KstConfig, settling_speed: float | None = None) -> KstConfig:
    """Derived parameters, in the reference's arithmetic (src/Parameters.f90:629-648)."""
    c.gred = (c.rhos / c.rhow - 1.0) * c.g
    c.rep = math.sqrt(c.g * c.solid_diameter) * c.solid_diameter / VISC_W
    R = (c.gred / VISC_W / VISC_W) ** (1.0 / 3.0) * c.solid_diameter
    if settling_speed is None:
        c.ws0 = VISC_W / c.solid_diameter * (math.sqrt(10.36 * 10.36 + 1.048 * R * R * R) - 10.36)
    else:
        c.ws0 = settling_speed
    c.nsettling = (4.7 + 0.41 * c.rep ** 0.75) / (1.0 + 0.175 * c.rep ** 0.75)
    c.critical_shields = 0.3 / (1.0 + 1.2 * R) + 0.055 * (1.0 - math.exp(-0.02 * R))
    c.diffusive_time_scale = 1.7976931348623157e308
    if c.eddy_viscosity > 0.0:
        c.diffusive_time_scale = min(c.dx * c.dx / c.eddy_viscosity, c.dy * c.dy / c.eddy_viscosity)
    return c


def set_sources(c: KstConfig, sources: list[FluxSource]) -> KstConfig:
    """Attach flux sources (num_cells must already be counted)."""
    arr = (KstFluxSource * max(1, len(sources)))()
    keep = [arr]
    for i, s in enumerate(sources):
        t = (ctypes.c_double * len(s.time))(*s.time)
        f = (ctypes.c_double * len(s.flux))(*s.flux)
        p = (ctypes.c_double * len(s.psi))(*s.psi)
        keep += [t, f, p]
        arr[i] = KstFluxSource(s.x, s.y, s.radius, len(s.time), int(s.num_cells),
                               ctypes.cast(t, ctypes.POINTER(ctypes.c_double)),
                               ctypes.cast(f, ctypes.POINTER(ctypes.c_double)),
                               ctypes.cast(p, ctypes.POINTER(ctypes.c_double)))
    c.n_sources = len(sources)
    c.sources = ctypes.cast(arr, ctypes.POINTER(KstFluxSource))
    c._keep = keep
    return c
