"""Scenario set-up on the host: topography, initial caps, flux sources.

This is the caller side of the time-step boundary -- the work Kestrel does
before ``Run`` in GetHeights (src/dem.f90:360-393), the TopogFunc family
(src/TopogFuncs.f90:82-603) and LoadSourceConditions (src/SetSources.f90:47-392)
-- restated with numpy in the reference's arithmetic so that the state handed
to the device matches the reference's tileContainer.

Arrays are global and row-major: cells (NY, NX), vertices (NY+1, NX+1), tile
state (ny_tiles, nx_tiles) flattened in TileID order.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .config import (BC_DIRICHLET, BC_HALT, BC_PERIODIC, BC_SPONGE, FluxSource, KstConfig,
                     TILE_ACTIVE, TILE_GHOST, TILE_OFF)


def cell_centres(cfg: KstConfig) -> tuple[np.ndarray, np.ndarray]:
    """Cell-centre coordinates of GridToPhysical (src/Grid.f90:339-353)."""
    nx, ny = cfg.nx_per_tile, cfg.ny_per_tile
    I = np.arange(cfg.NX)
    J = np.arange(cfg.NY)
    gi, ii = I // nx, I % nx
    gj, jj = J // ny, J % ny
    x = -0.5 * cfg.x_size + cfg.dx * ((gi + 1 - 1.0) * nx + (ii + 1 - 0.5))
    y = -0.5 * cfg.y_size + cfg.dy * ((gj + 1 - 1.0) * ny + (jj + 1 - 0.5))
    return x.astype(np.float64), y.astype(np.float64)


def vertex_coords(cfg: KstConfig) -> tuple[np.ndarray, np.ndarray]:
    """Vertex coordinates (src/UpdateTiles.f90:312-327) after the tile-edge
    equalisation convention (the east/north tile's value wins)."""
    x, y = cell_centres(cfg)
    xv = np.empty(cfg.NX + 1)
    yv = np.empty(cfg.NY + 1)
    xv[:-1] = x - 0.5 * cfg.dx
    xv[-1] = x[-1] + 0.5 * cfg.dx
    yv[:-1] = y - 0.5 * cfg.dy
    yv[-1] = y[-1] + 0.5 * cfg.dy
    return xv, yv


def topo_xslope(cfg: KstConfig, slope: float) -> np.ndarray:
    """xslope (src/TopogFuncs.f90:404-418): b0 = slope * x at the vertices."""
    xv, yv = vertex_coords(cfg)
    return np.ascontiguousarray(np.broadcast_to(slope * xv[None, :], (cfg.NY + 1, cfg.NX + 1)))


def topo_flat(cfg: KstConfig, level: float = 0.0) -> np.ndarray:
    return np.full((cfg.NY + 1, cfg.NX + 1), float(level))


@dataclass
class Hills:
    """Seeded sum of Gaussian hills on an x-ramp (a TopogFunc of our own; the
    reference has no such function)."""
    seed: int
    n: int = 32
    amp: tuple = (5.0, 50.0)
    width: tuple = (50.0, 300.0)
    ramp_deg: float = 5.0
    centres: np.ndarray = field(default=None, repr=False)
    amps: np.ndarray = field(default=None, repr=False)
    widths: np.ndarray = field(default=None, repr=False)


def topo_gaussian_hills(cfg: KstConfig, hills: Hills) -> np.ndarray:
    rng = np.random.default_rng(hills.seed)
    hills.amps = rng.uniform(hills.amp[0], hills.amp[1], hills.n)
    hills.widths = rng.uniform(hills.width[0], hills.width[1], hills.n)
    cx = rng.uniform(-0.5 * cfg.x_size, 0.5 * cfg.x_size, hills.n)
    cy = rng.uniform(-0.5 * cfg.y_size, 0.5 * cfg.y_size, hills.n)
    hills.centres = np.stack([cx, cy], axis=1)
    xv, yv = vertex_coords(cfg)
    X, Y = np.meshgrid(xv, yv)
    b = math.tan(math.radians(hills.ramp_deg)) * X
    for a, s, x0, y0 in zip(hills.amps, hills.widths, cx, cy):
        b = b + a * np.exp(-((X - x0) ** 2 + (Y - y0) ** 2) / (2.0 * s * s))
    return np.ascontiguousarray(b)


def _kahan(*xs):
    """KahanSum (src/utilities.f90:418-448), elementwise over arrays."""
    s = np.zeros_like(xs[0])
    c = np.zeros_like(xs[0])
    for x in xs:
        y = x - c
        t = s + y
        c = (t - s) - y
        s = t
    return s


def is_1d(cfg: KstConfig) -> bool:
    """isOneD (src/DomainSettings.f90:216-220): one row of cells."""
    return cfg.ny_tiles * cfg.ny_per_tile == 1


def cell_topography(cfg: KstConfig, b0v: np.ndarray, btv: np.ndarray | None = None):
    """Cell-centred b0, bt, db/dx, db/dy
    (ComputeCellCentredTopographicData, src/MorphodynamicRHS.f90:308-360; on a
    1D grid two-vertex forms along the lower vertex row, db/dy zero)."""
    if btv is None:
        btv = np.zeros_like(b0v)
    if is_1d(cfg):
        b, t = b0v[0], btv[0]
        b0c = 0.5 * (b[:-1] + b[1:])
        btc = 0.5 * (t[:-1] + t[1:])
        dbdx = (1.0 / cfg.dx) * _kahan(b[1:], t[1:], -b[:-1], -t[:-1])
        return b0c[None, :], btc[None, :], dbdx[None, :], np.zeros((1, cfg.NX))
    b00, b10, b01, b11 = b0v[:-1, :-1], b0v[:-1, 1:], b0v[1:, :-1], b0v[1:, 1:]
    t00, t10, t01, t11 = btv[:-1, :-1], btv[:-1, 1:], btv[1:, :-1], btv[1:, 1:]
    b0c = 0.25 * _kahan(b00, b10, b01, b11)
    btc = 0.25 * _kahan(t00, t10, t01, t11)
    dxr, dyr = 1.0 / cfg.dx, 1.0 / cfg.dy
    dbdx = 0.5 * dxr * _kahan(b10, t10, -b00, -t00, b11, t11, -b01, -t01)
    dbdy = 0.5 * dyr * _kahan(b01, t01, -b00, -t00, b11, t11, -b10, -t10)
    return b0c, btc, dbdx, dbdy


def gamma(cfg: KstConfig, bx, by):
    if not cfg.geometric_factors:
        return np.ones_like(bx)
    return np.sqrt(1.0 + bx * bx + by * by)


@dataclass
class Cap:
    x: float
    y: float
    radius: float
    height: float
    psi: float = 0.0
    u: float = 0.0
    v: float = 0.0
    shape: str = "para"    # 'flat' | 'para' | 'level'

    @staticmethod
    def from_volume(x, y, radius, volume, shape="para", **kw):
        """capVolume -> capHeight (src/InitConds.f90:362-368): a paraboloid of
        radius R and height H holds pi R^2 H / 2, a flat cap pi R^2 H."""
        if shape == "flat":
            h = volume / math.pi / radius / radius
        else:
            h = 2.0 * volume / math.pi / radius / radius
        return Cap(x, y, radius, h, shape=shape, **kw)


@dataclass
class Cube:
    """A cuboid initial condition (Cube_Set, src/InitConds.f90:523-660): cells
    whose centres lie within length/2 and width/2 of (x, y) get a flat layer
    of depth `height` or, with shape 'level', a free surface at `height`; its
    velocities are not painted (src/SetSources.f90:305-355)."""
    x: float
    y: float
    length: float
    width: float
    height: float
    psi: float = 0.0
    shape: str = "flat"    # 'flat' | 'level'


@dataclass
class InitialState:
    w: np.ndarray
    rhoHnu: np.ndarray
    rhoHnv: np.ndarray
    Hnpsi: np.ndarray
    tile_state: np.ndarray
    hnmax0: np.ndarray
    # the running maximum of the solids fraction a 1D cap paints
    # (src/SetSources.f90:131-160); None where nothing is painted
    psimax0: np.ndarray | None = None


def _ghost_ring(cfg: KstConfig, active: np.ndarray) -> np.ndarray:
    """Tile states: active tiles plus their 8 neighbours as ghosts
    (AddGhostTiles, src/UpdateTiles.f90:389-481); with periodic boundaries the
    neighbours wrap around the domain and there are no corner ghosts."""
    a = active.reshape(cfg.ny_tiles, cfg.nx_tiles)
    g = np.zeros_like(a)
    periodic = cfg.bc == BC_PERIODIC
    for dj in (-1, 0, 1):
        for di in (-1, 0, 1):
            if periodic and di != 0 and dj != 0:
                continue
            if is_1d(cfg) and dj != 0:   # W and E only on a 1D grid
                continue
            if periodic:
                sh = np.roll(np.roll(a, dj, axis=0), di, axis=1)
            else:
                sh = np.zeros_like(a)
                ys = slice(max(dj, 0), cfg.ny_tiles + min(dj, 0))
                yd = slice(max(-dj, 0), cfg.ny_tiles + min(-dj, 0))
                xs = slice(max(di, 0), cfg.nx_tiles + min(di, 0))
                xd = slice(max(-di, 0), cfg.nx_tiles + min(-di, 0))
                sh[yd, xd] = a[ys, xs]
            g |= sh
    st = np.full(a.shape, TILE_OFF, dtype=np.uint8)
    st[g.astype(bool)] = TILE_GHOST
    st[a.astype(bool)] = TILE_ACTIVE
    return st.reshape(-1)


def _tile_of_cells(cfg: KstConfig) -> np.ndarray:
    J, I = np.meshgrid(np.arange(cfg.NY), np.arange(cfg.NX), indexing="ij")
    return (J // cfg.ny_per_tile) * cfg.nx_tiles + (I // cfg.nx_per_tile)


def _edge_tiles(cfg: KstConfig) -> np.ndarray:
    """OnDomainEdge (src/Grid.f90:322-335) per tile."""
    edge = np.zeros((cfg.ny_tiles, cfg.nx_tiles), dtype=bool)
    edge[:, 0] = edge[:, -1] = True
    if cfg.ny_tiles > 1:
        edge[0, :] = edge[-1, :] = True
    return edge.reshape(-1)


def _check_halt(cfg: KstConfig, active: np.ndarray):
    if cfg.bc == BC_HALT and np.any(active.astype(bool) & _edge_tiles(cfg)):
        raise ValueError("Error: tried to add a tile outside the domain "
                         "(halt boundary; src/UpdateTiles.f90:61-65)")


def initial_state(cfg: KstConfig, b0v: np.ndarray, caps: list[Cap] = (),
                  sources: list[FluxSource] = (), cubes: list[Cube] = ()) -> InitialState:
    """Paint caps and cubes and count flux-source cells as LoadSourceConditions
    does (src/SetSources.f90:228-303 caps, 305-355 cubes, 357-375 flux
    sources; on a 1D grid :95-240: distances along x only, no y velocity).
    Only a tile that AddTile activates is painted: with dirichlet or sponge
    boundaries a tile on the domain edge is silently left inactive
    (src/UpdateTiles.f90:61-69)."""
    NX, NY = cfg.NX, cfg.NY
    oned = is_1d(cfg)
    x, y = cell_centres(cfg)
    X, Y = np.meshgrid(x, y)
    if oned:
        Y = None
    b0c, btc, bx, by = cell_topography(cfg, b0v)
    gam = gamma(cfg, bx, by)
    tiles = _tile_of_cells(cfg)
    active = np.zeros(cfg.ntiles, dtype=np.uint8)
    w = b0c.copy()   # ActivateTile sets w = b0 (cell centred), src/UpdateTiles.f90:363-366
    rhu = np.zeros((NY, NX))
    rhv = np.zeros((NY, NX))
    hnpsi = np.zeros((NY, NX))
    hn = np.zeros((NY, NX))
    psimax = np.zeros((NY, NX))
    # tiles AddTile may activate (the others stay off and are not painted)
    allowed = np.ones(cfg.ntiles, dtype=bool)
    if cfg.bc in (BC_DIRICHLET, BC_SPONGE):
        allowed &= ~_edge_tiles(cfg)
    cell_ok = allowed[tiles]
    def dist2(px, py):
        if oned:
            return (X - px) * (X - px)
        return (X - px) * (X - px) + (Y - py) * (Y - py)

    for cap in caps:
        R2 = dist2(cap.x, cap.y)
        inside = (R2 <= cap.radius * cap.radius)
        active[np.unique(tiles[inside & cell_ok])] = 1
        inside = inside & cell_ok
        rho = cfg.rhow + (cfg.rhos - cfg.rhow) * cap.psi
        cv = 0.0 if oned else cap.v
        if cap.shape == "flat":
            inc_w = cap.height / gam
            inc_h = np.full_like(gam, cap.height)
            inc_rhu = np.full_like(gam, rho * cap.height * cap.u)
            inc_rhv = np.full_like(gam, rho * cap.height * cv)
            inc_psi = np.full_like(gam, cap.psi * cap.height)
        elif cap.shape == "para":
            prof = 1.0 - R2 / cap.radius / cap.radius
            inc_w = cap.height * prof / gam
            inc_h = cap.height * prof
            inc_rhu = np.full_like(gam, rho * cap.height * cap.u)
            inc_rhv = np.full_like(gam, rho * cap.height * cv)
            inc_psi = cap.psi * cap.height * prof
        elif cap.shape == "level":
            hp = cap.height - b0c
            Hn = hp * gam
            pos = Hn > 0.0
            inside = inside & pos
            inc_w = hp
            inc_h = Hn
            inc_rhu = rho * Hn * cap.u
            inc_rhv = rho * Hn * cv
            inc_psi = cap.psi * Hn
        else:
            raise ValueError(f"unknown cap shape {cap.shape!r}")
        w = np.where(inside, w + inc_w, w)
        hn = np.where(inside, hn + inc_h, hn)
        rhu = np.where(inside, rhu + inc_rhu, rhu)
        rhv = np.where(inside, rhv + inc_rhv, rhv)
        hnpsi = np.where(inside, hnpsi + inc_psi, hnpsi)
        if oned:
            psimax = np.where(inside, psimax + cap.psi, psimax)
    for cube in cubes:
        inside = np.abs(X - cube.x) <= 0.5 * cube.length
        if not oned:
            inside = inside & (np.abs(Y - cube.y) <= 0.5 * cube.width)
        active[np.unique(tiles[inside & cell_ok])] = 1
        inside = inside & cell_ok
        if cube.shape == "level":
            hp = cube.height - b0c
            Hn = hp * gam
            inside = inside & (Hn > 0.0)
            inc_w, inc_h, inc_psi = hp, Hn, Hn * cube.psi
        else:
            inc_w = cube.height / gam
            inc_h = np.full_like(gam, cube.height)
            inc_psi = np.full_like(gam, cube.psi * cube.height)
        w = np.where(inside, w + inc_w, w)
        hn = np.where(inside, hn + inc_h, hn)
        rhu = np.where(inside, rhu + 0.0, rhu)
        rhv = np.where(inside, rhv + 0.0, rhv)
        hnpsi = np.where(inside, hnpsi + inc_psi, hnpsi)
    for src in sources:
        R2 = dist2(src.x, src.y)
        inside = R2 <= src.radius * src.radius
        src.num_cells = int(np.count_nonzero(inside))
        active[np.unique(tiles[inside & cell_ok])] = 1
    _check_halt(cfg, active)
    state = _ghost_ring(cfg, active)
    return InitialState(np.ascontiguousarray(w), rhu, rhv, hnpsi, state, hn,
                        psimax if oned and caps else None)
