"""Restart from the reference's text output files.

LoadInitialCondition_txt (src/Restart.f90:151-328) reads back a solution file
(OutputSolutionData_txt's layout: a header, then per listed tile nYpertile
rows of nXpertile lines `tile, x, y, lat, lon, Hn, speed, u, v, rho, psi,
rhoHnu, rhoHnv, Hnpsi, w, b0, bt, dbdx, dbdy`, an empty record after each
row), its `_topo` companion (per tile its (nX+1) x (nY+1) vertices `tile, b0,
bt`) and, on a restart, the maxima file.  The listed tiles become the active
tiles; their rhoHnu, rhoHnv, Hnpsi and w, their vertex bt and the maxima are
taken from the files.

The values carry the files' 11 significant digits.  One difference from the
reference: it also reads each cell's bt, db/dx and db/dy from the solution
file, while kst_set_state derives them from the vertex bt (as
ComputeCellCentredTopographicData does during a run), so the restarted cell
topography may differ from the reference's in the 11th digit.  Each tile's
own copy of a shared vertex is read, and the copy of the tile holding the
vertex's cell (the one kst_get_state reports) is the one given to
kst_set_state.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import scenario
from .config import KstConfig


def _records(path: str):
    """The data lines of a file after its header: (tile id, values)."""
    with open(path) as f:
        next(f)
        for line in f:
            if not line.strip():
                continue
            parts = line.split(",")
            yield int(parts[0]), [float(x) for x in parts[1:]]


def read_solution_txt(cfg: KstConfig, path: str):
    """(tiles, cols): the listed tiles (0-based ids, file order) and the 18
    per-cell columns after the tile number as global NY x NX arrays (zero
    outside the listed tiles)."""
    nx, ny, ntx = cfg.nx_per_tile, cfg.ny_per_tile, cfg.nx_tiles
    cols = np.zeros((18, cfg.NY, cfg.NX))
    tiles: list[int] = []
    n = 0
    for ttk, vals in _records(path):
        if len(vals) != 18:
            raise ValueError(f"{path}: expected 18 values after the tile number, got {len(vals)}")
        q, c = divmod(n, nx * ny)
        k = ttk - 1
        if c == 0:
            tiles.append(k)
        elif tiles[q] != k:
            raise ValueError(f"{path}: tile {ttk} interrupts the block of tile {tiles[q] + 1}")
        j, i = divmod(c, nx)
        gi, gj = k % ntx, k // ntx
        cols[:, gj * ny + j, gi * nx + i] = vals
        n += 1
    if n % (nx * ny):
        raise ValueError(f"{path}: incomplete tile block ({n} cell lines)")
    return tiles, cols


def read_topo_txt(cfg: KstConfig, path: str) -> np.ndarray:
    """Each listed tile's own vertex bt, [ntiles, ny+1, nx+1] (zero for
    tiles not listed)."""
    nx, ny = cfg.nx_per_tile, cfg.ny_per_tile
    nv = (nx + 1) * (ny + 1)
    out = np.zeros((cfg.ntiles, ny + 1, nx + 1))
    n = 0
    for ttk, vals in _records(path):
        if len(vals) != 2:
            raise ValueError(f"{path}: expected b0, bt after the tile number")
        c = n % nv
        j, i = divmod(c, nx + 1)
        out[ttk - 1, j, i] = vals[1]
        n += 1
    if n % nv:
        raise ValueError(f"{path}: incomplete tile block ({n} vertex lines)")
    return out


def read_maxima_txt(cfg: KstConfig, path: str) -> list[np.ndarray]:
    """The 11 maxima arrays (kst_get_maxima's order) as global NY x NX arrays."""
    nx, ny, ntx = cfg.nx_per_tile, cfg.ny_per_tile, cfg.nx_tiles
    mx = [np.zeros((cfg.NY, cfg.NX)) for _ in range(11)]
    n = 0
    for ttk, vals in _records(path):
        if len(vals) != 15:
            raise ValueError(f"{path}: expected 15 values after the tile number")
        c = n % (nx * ny)
        j, i = divmod(c, nx)
        k = ttk - 1
        J, I = (k // ntx) * ny + j, (k % ntx) * nx + i
        for q in range(11):
            mx[q][J, I] = vals[4 + q]
        n += 1
    return mx


def global_vertex_bt(cfg: KstConfig, tile_bt: np.ndarray, tiles) -> np.ndarray:
    """Vertex (I, J) from the copy of the tile holding cell
    (min(I, NX-1), min(J, NY-1)) when that tile is listed, else from a listed
    tile that shares the vertex; 0 where none does."""
    nx, ny, ntx = cfg.nx_per_tile, cfg.ny_per_tile, cfg.nx_tiles
    listed = np.zeros(cfg.ntiles, dtype=bool)
    listed[list(tiles)] = True
    btv = np.zeros((cfg.NY + 1, cfg.NX + 1))
    for J in range(cfg.NY + 1):
        for I in range(cfg.NX + 1):
            ci, cj = min(I, cfg.NX - 1), min(J, cfg.NY - 1)
            for di, dj in ((0, 0), (-1, 0), (0, -1), (-1, -1)):
                ti, tj = (ci // nx) + di, (cj // ny) + dj
                if ti < 0 or tj < 0:
                    continue
                k = tj * ntx + ti
                if not listed[k]:
                    continue
                li, lj = I - ti * nx, J - tj * ny
                if 0 <= li <= nx and 0 <= lj <= ny:
                    btv[J, I] = tile_bt[k, lj, li]
                    break
    return btv


@dataclass
class RestartState:
    init: scenario.InitialState
    btv: np.ndarray
    maxima: list | None
    tiles: list
    mask: np.ndarray          # cells of the listed tiles


def load_txt(cfg: KstConfig, b0v: np.ndarray, result_path: str, topo_path: str | None = None,
             maxima_path: str | None = None) -> RestartState:
    """Read a solution file, its `_topo` companion (default result_path +
    '_topo') and optionally a maxima file into the arguments of
    kst_set_state (and kst_set_maxima)."""
    tiles, cols = read_solution_txt(cfg, result_path)
    tile_bt = read_topo_txt(cfg, topo_path or result_path + "_topo")
    active = np.zeros(cfg.ntiles, dtype=np.uint8)
    active[tiles] = 1
    # cells outside the listed tiles as a fresh tile holds them (w = b0)
    base = scenario.initial_state(cfg, b0v, [])
    mask = np.zeros((cfg.NY, cfg.NX), dtype=bool)
    nx, ny, ntx = cfg.nx_per_tile, cfg.ny_per_tile, cfg.nx_tiles
    for k in tiles:
        gi, gj = k % ntx, k // ntx
        mask[gj * ny:(gj + 1) * ny, gi * nx:(gi + 1) * nx] = True
    # solution file columns (after the tile number): 10 rhoHnu, 11 rhoHnv, 12 Hnpsi, 13 w
    w = np.where(mask, cols[13], base.w)
    rhu = np.where(mask, cols[10], 0.0)
    rhv = np.where(mask, cols[11], 0.0)
    hnpsi = np.where(mask, cols[12], 0.0)
    tile_state = scenario._ghost_ring(cfg, active)
    init = scenario.InitialState(w, rhu, rhv, hnpsi, tile_state, None)
    btv = global_vertex_bt(cfg, tile_bt, tiles)
    mx = read_maxima_txt(cfg, maxima_path) if maxima_path else None
    return RestartState(init, btv, mx, tiles, mask)


def restart(stepper, cfg: KstConfig, b0v: np.ndarray, t: float, result_path: str,
            topo_path: str | None = None, maxima_path: str | None = None) -> RestartState:
    """Set a Stepper's state (and maxima) from text output written at time t.
    The maxima of cells outside the listed tiles keep the values kst_set_state
    starts them with (a tile activated later starts its maxima afresh)."""
    rs = load_txt(cfg, b0v, result_path, topo_path, maxima_path)
    stepper.set_state(rs.init, t=t, btv=rs.btv)
    if rs.maxima is not None:
        cur = stepper.maxima()
        for q in range(11):
            cur[q][rs.mask] = rs.maxima[q][rs.mask]
        stepper.set_maxima(cur)
    return rs
