"""Restart from the reference's text output files.

LoadInitialCondition_txt (src/Restart.f90:151-328) reads back a solution file
(OutputSolutionData_txt's layout: a header, then per listed tile nYpertile
rows of nXpertile lines `tile, x, y, lat, lon, Hn, speed, u, v, rho, psi,
rhoHnu, rhoHnv, Hnpsi, w, b0, bt, dbdx, dbdy`, an empty record after each
row), its `_topo` companion (per tile its (nX+1) x (nY+1) vertices `tile, b0,
bt`) and, on a restart, the maxima file.  The listed tiles become the active
tiles; their rhoHnu, rhoHnv, Hnpsi and w, their vertex bt and the maxima are
taken from the files.

The values carry the files' 11 significant digits.  As in the reference,
the listed cells' bt, db/dx and db/dy are read from the solution file and kept
(kst_set_cell_topography) where the reference keeps them:
ActivateTile(restart=.true.) rebuilds only the face topography from the
vertices (src/UpdateTiles.f90:342-384), but the ghost tiles it adds derive the
cell topography of some listed neighbours again from the vertices
(file_topography_tiles), and the morphodynamic operator derives it again at
its first stage.  Each tile's own copy of a shared vertex is read, and the copy of the
tile holding the vertex's cell (the one kst_get_state reports) is the one
given to kst_set_state.  Cell coordinates are the grid's own, which equal the
file's where they are exactly representable (every deck here).

A restart of a run (Restart = on, src/Restart.f90:55-146) reads RunInfo.txt
(ReadRunInfoToRunParams, src/Restart.f90:560-638): the time step between
outputs, the last output file and t start; it resumes at t start + that
step x the last file's number from the last file (restart_run).  Text files
of 2D grids only (a 1D grid's solution file has other columns).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import scenario
from .config import BC_PERIODIC, KstConfig


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


def _ref_neighbour(cfg: KstConfig, k: int, di: int, dj: int) -> int:
    """Tile k's neighbour as the reference's tile fields hold it
    (SetTileBoundaries, src/UpdateTiles.f90:486-517): ids offset by 1 and by
    nXtiles (so a West across the domain edge is the previous row's last
    tile), wrapped with periodic boundaries (SetPeriodicBoundaryConds,
    src/UpdateTiles.f90:521-566); -1 outside the tile range."""
    ntx, nty = cfg.nx_tiles, cfg.ny_tiles
    if cfg.bc == BC_PERIODIC:
        i, j = (k % ntx + di) % ntx, (k // ntx + dj) % nty
        return j * ntx + i
    t = k + di + dj * ntx
    return t if 0 <= t < cfg.ntiles else -1


def file_topography_tiles(cfg: KstConfig, tiles) -> set:
    """The listed tiles whose cell bt, db/dx, db/dy a restart keeps from the
    solution file.  ActivateTile(restart=.true.) of each listed tile in the
    active list's order (src/Restart.f90:286-289; ascending ids, as
    AddToActiveTiles keeps the list, src/UpdateTiles.f90:81-117) adds its new ghost tiles
    (AddGhostTiles, src/UpdateTiles.f90:389-481: W, E, N, S, then SW, SE, NW,
    NE but with periodic boundaries, W and E on a 1D grid; a tile neither
    listed nor a ghost already), and EqualiseTopographicBoundaryData of each
    new ghost G (src/MorphodynamicRHS.f90:588-687) derives again from the
    vertices the cell topography of W(G), of S(W(G)), of S(G) and, when
    S(W(G)) was not derived, of W(S(G)) -- each when it is listed or a ghost."""
    oned = cfg.ny_tiles * cfg.ny_per_tile == 1
    periodic = cfg.bc == BC_PERIODIC
    listed = sorted(int(k) for k in tiles)
    lset = set(listed)
    dirs = [(-1, 0), (1, 0)] if oned else [(-1, 0), (1, 0), (0, 1), (0, -1)]
    if not oned and not periodic:
        dirs += [(-1, -1), (1, -1), (-1, 1), (1, 1)]
    ghosts: set = set()
    derived: set = set()
    for k in listed:
        new: list = []
        for di, dj in dirs:
            t = _ref_neighbour(cfg, k, di, dj)
            if t < 0 or t in ghosts or t in lset or t in new:
                continue
            new.append(t)
        ghosts.update(new)

        def on(x):
            return x >= 0 and (x in lset or x in ghosts)

        for g in new:
            sw_done = False
            w = _ref_neighbour(cfg, g, -1, 0)
            if on(w):
                derived.add(w)
                if not oned:
                    sw = _ref_neighbour(cfg, w, 0, -1)
                    if on(sw):
                        derived.add(sw)
                        sw_done = True
            if not oned:
                sth = _ref_neighbour(cfg, g, 0, -1)
                if on(sth):
                    derived.add(sth)
                    if not sw_done:
                        sw = _ref_neighbour(cfg, sth, -1, 0)
                        if on(sw):
                            derived.add(sw)
    return lset - derived


@dataclass
class RestartState:
    init: scenario.InitialState
    btv: np.ndarray
    maxima: list | None
    tiles: list
    mask: np.ndarray          # cells of the listed tiles
    btc: np.ndarray           # the file's cell bt, db/dx, db/dy (NY x NX)
    dbdx: np.ndarray
    dbdy: np.ndarray
    keep: np.ndarray          # cells whose file topography the restart keeps


def load_txt(cfg: KstConfig, b0v: np.ndarray, result_path: str, topo_path: str | None = None,
             maxima_path: str | None = None) -> RestartState:
    """Read a solution file, its `_topo` companion (default result_path +
    '_topo') and optionally a maxima file into the arguments of
    kst_set_state (and kst_set_maxima)."""
    if cfg.ny_tiles * cfg.ny_per_tile == 1:
        raise ValueError("text restart: 2D grids only")
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
    keep = np.zeros((cfg.NY, cfg.NX), dtype=bool)
    for k in file_topography_tiles(cfg, tiles):
        gi, gj = k % ntx, k // ntx
        keep[gj * ny:(gj + 1) * ny, gi * nx:(gi + 1) * nx] = True
    # columns 15-17: the cells' bt, db/dx, db/dy
    return RestartState(init, btv, mx, tiles, mask, cols[15], cols[16], cols[17], keep)


def restart(stepper, cfg: KstConfig, b0v: np.ndarray, t: float, result_path: str,
            topo_path: str | None = None, maxima_path: str | None = None) -> RestartState:
    """Set a Stepper's state (and maxima) from text output written at time t.
    The maxima of cells outside the listed tiles keep the values kst_set_state
    starts them with (a tile activated later starts its maxima afresh)."""
    rs = load_txt(cfg, b0v, result_path, topo_path, maxima_path)
    stepper.set_state(rs.init, t=t, btv=rs.btv)
    stepper.set_cell_topography(rs.btc, rs.dbdx, rs.dbdy, rs.keep)
    if rs.maxima is not None:
        cur = stepper.maxima()
        for q in range(11):
            cur[q][rs.mask] = rs.maxima[q][rs.mask]
        stepper.set_maxima(cur)
    return rs


@dataclass
class RunInfo:
    delta_t: float            # time step between outputs
    last_file: int            # number of the last output file
    tstart: float


def read_run_info(path: str) -> RunInfo:
    """ReadRunInfoToRunParams (src/Restart.f90:560-638): the lines of the form
    label = value whose lower-cased label contains 'time step between
    outputs', 'last output file' or 't start'; each must be present."""
    found: dict = {}
    with open(path, errors="replace") as f:
        for line in f:
            line = line.strip()
            if "=" not in line:
                continue
            label, val = line.split("=", 1)
            label = label.lower()
            if "time step between outputs" in label:
                found["delta_t"] = float(val)
            if "last output file" in label:
                found["last_file"] = int(val)
            if "t start" in label:
                found["tstart"] = float(val)
    for k, what in (("delta_t", "Time step between outputs"), ("last_file", "Last output file"),
                    ("tstart", "t start")):
        if k not in found:
            raise ValueError(f"Could not find '{what}' in {path}")
    return RunInfo(**found)


def restart_run(stepper, cfg: KstConfig, b0v: np.ndarray, out_dir: str) -> tuple:
    """Resume a run whose output folder out_dir holds RunInfo.txt, the output
    files and Maximums.txt, as LoadInitialCondition does with Restart = on
    (src/Restart.f90:55-146): t = t start + time step between outputs x the
    last file's number (grid%t, src/Restart.f90:79), state from that file
    (GenerateOutputFilename's six-digit name, src/Restart.f90:553), its _topo
    companion and the maxima.  Returns (RestartState, RunInfo, t)."""
    import os
    info = read_run_info(os.path.join(out_dir, "RunInfo.txt"))
    t = info.tstart + info.delta_t * info.last_file
    path = os.path.join(out_dir, f"{info.last_file:06d}.txt")
    rs = restart(stepper, cfg, b0v, t, path, path + "_topo", os.path.join(out_dir, "Maximums.txt"))
    return rs, info, t
