"""Output files of the reference, written from device results.

CalculateVolume (src/Output.f90:617-735) appends one line per output time to
Volume.txt: the time as F12.2, then six ES24.15 columns separated by ', ',
under a header of right-justified names (src/Output.f90:664-673, 720-727).
The six values come from the device (`Stepper.volume`, i.e. kst_volume).

OutputSolutionData_txt (src/Output.f90:754-929) writes, for the active tiles
in list order and their cells row by row, the tile number (I8) and 19
ES18.10E3 columns, with an empty record after each row of a tile; and a
`_topo` file of the tiles' vertex b0 and bt.  The per-cell columns come from
the device (`Stepper.output_fields`, i.e. kst_get_output); the coordinates
are GridToPhysical's (src/Grid.f90:339-353); latitude and longitude are 0
(no georeferencing).

OutputMaximums_txt (src/Output.f90:1270-1359) writes the running maxima the
same way, with an empty record after each row and after each tile; they come
from the device (`Stepper.maxima`, i.e. kst_get_maxima).
"""
from __future__ import annotations

import math
import os

import numpy as np

from . import scenario

VOLUME_COLUMNS = ("time", "volume", "total_bed_volume", "total_mass", "bed_mass",
                  "total_solids_mass", "bed_solids_mass")


def fortran_f(x: float, w: int, d: int) -> str:
    """Fortran Fw.d edit descriptor."""
    if math.isnan(x):
        s = "NaN"
    elif math.isinf(x):
        s = "Infinity" if x > 0 else "-Infinity"
    else:
        s = f"{x:.{d}f}"
    return "*" * w if len(s) > w else s.rjust(w)


def fortran_es(x: float, w: int, d: int, e: int | None = None) -> str:
    """Fortran ESw.d (e is None) or ESw.dEe edit descriptor.  Without Ee the
    exponent takes two digits after 'E', or three with the 'E' dropped when
    |exponent| > 99; with Ee it always has e digits after 'E'."""
    if math.isnan(x):
        s = "NaN"
    elif math.isinf(x):
        s = "Infinity" if x > 0 else "-Infinity"
    else:
        mant, ex = f"{x:.{d}E}".split("E")
        n = int(ex)
        sign = "+" if n >= 0 else "-"
        if e is not None:
            digits = f"{abs(n):0{e}d}"
            s = "*" * w if len(digits) > e else f"{mant}E{sign}{digits}"
        elif abs(n) <= 99:
            s = f"{mant}E{sign}{abs(n):02d}"
        else:
            s = f"{mant}{sign}{abs(n):03d}"
    return "*" * w if len(s) > w else s.rjust(w)


def volume_header() -> str:
    """The header line of Volume.txt (format (a12,6(a2,a24)))."""
    return VOLUME_COLUMNS[0].rjust(12) + "".join(", " + c.rjust(24) for c in VOLUME_COLUMNS[1:])


def volume_line(t: float, cols) -> str:
    """One Volume.txt line (format (f12.2,6(a2, es24.15)))."""
    if len(cols) != 6:
        raise ValueError("Volume.txt holds six values per line")
    return fortran_f(t, 12, 2) + "".join(", " + fortran_es(float(v), 24, 15) for v in cols)


def write_volume(path: str, t: float, cols, create: bool = False) -> None:
    """Append a line to Volume.txt at path, starting the file (with its header)
    when it does not exist or create is set, as CalculateVolume does."""
    new = create or not os.path.exists(path)
    with open(path, "w" if new else "a") as f:
        if new:
            f.write(volume_header() + "\n")
        f.write(volume_line(t, cols) + "\n")


SOLUTION_COLUMNS = ("tile", "x", "y", "latitude", "longitude", "flow_depth", "flow_speed",
                    "x_velocity", "y_velocity", "density", "solids_fraction", "x_flux", "y_flux",
                    "Hnpsi", "w", "base_elevation", "elevation_change", "x_slope", "y_slope")
TOPO_COLUMNS = ("tile", "B0_vertex", "Bt_vertex")


def solution_header() -> str:
    """Header of the 2D solution file (format (a8,a2,18(a18,a2),a18))."""
    return SOLUTION_COLUMNS[0].rjust(8) + "".join(", " + c.rjust(18) for c in SOLUTION_COLUMNS[1:])


def topo_header() -> str:
    """Header of the _topo file (format (a8, a2, a18, a2, a18))."""
    return TOPO_COLUMNS[0].rjust(8) + "".join(", " + c.rjust(18) for c in TOPO_COLUMNS[1:])


def _row(tile_id: int, vals) -> str:
    return f"{tile_id:8d}" + "".join(", " + fortran_es(float(v), 18, 10, 3) for v in vals)


def solution_lines(cfg, tiles, fields):
    """The lines of the 2D solution file for tiles (0-based ids, list order)
    and fields[q, 14, ny_per_tile, nx_per_tile] (columns 6-19)."""
    xc, yc = scenario.cell_centres(cfg)
    nx, ny, ntx = cfg.nx_per_tile, cfg.ny_per_tile, cfg.nx_tiles
    out = [solution_header()]
    for q, k in enumerate(tiles):
        k = int(k)
        gi, gj = k % ntx, k // ntx
        for j in range(ny):
            y = yc[gj * ny + j]
            for i in range(nx):
                out.append(_row(k + 1, [xc[gi * nx + i], y, 0.0, 0.0, *fields[q, :, j, i]]))
            out.append("")
    return out


def topo_lines(cfg, tiles, b0v: np.ndarray, tile_bt: np.ndarray):
    """The lines of the _topo file: each listed tile's vertices, row by row,
    with its b0 (from the global vertex b0) and its own copy of bt
    (tile_bt[k, ny+1, nx+1])."""
    nx, ny, ntx = cfg.nx_per_tile, cfg.ny_per_tile, cfg.nx_tiles
    out = [topo_header()]
    for k in tiles:
        k = int(k)
        gi, gj = k % ntx, k // ntx
        for j in range(ny + 1):
            for i in range(nx + 1):
                out.append(_row(k + 1, [b0v[gj * ny + j, gi * nx + i], tile_bt[k, j, i]]))
            out.append("")
    return out


def write_solution_txt(path: str, cfg, tiles, fields, b0v: np.ndarray | None = None,
                       tile_bt: np.ndarray | None = None) -> None:
    """Write the solution file at path and, given b0v and tile_bt, its
    path + '_topo' companion (both replaced if present)."""
    with open(path, "w") as f:
        f.write("\n".join(solution_lines(cfg, tiles, fields)) + "\n")
    if b0v is not None and tile_bt is not None:
        with open(path + "_topo", "w") as f:
            f.write("\n".join(topo_lines(cfg, tiles, b0v, tile_bt)) + "\n")


MAXIMA_COLUMNS = ("tile", "x_distance", "y_distance", "latitude", "longitude", "max_depth",
                  "t_max_depth", "max_speed", "t_max_speed", "max_erosion", "t_max_erosion",
                  "max_deposit", "t_max_deposit", "max_solids_frac", "t_max_solids_frac",
                  "inundation_time")


def maxima_header() -> str:
    """Header of the maxima file (format (a8,a2,14(a18,a2), a18))."""
    return MAXIMA_COLUMNS[0].rjust(8) + "".join(", " + c.rjust(18) for c in MAXIMA_COLUMNS[1:])


def maxima_lines(cfg, tiles, maxima):
    """The lines of the maxima file for tiles (0-based ids, list order) and
    the 11 global NY x NX arrays of kst_get_maxima (Hnmax, its time, umax, its
    time, emax, its time, dmax, its time, psimax, its time, tfirst)."""
    xc, yc = scenario.cell_centres(cfg)
    nx, ny, ntx = cfg.nx_per_tile, cfg.ny_per_tile, cfg.nx_tiles
    out = [maxima_header()]
    for k in tiles:
        k = int(k)
        gi, gj = k % ntx, k // ntx
        for j in range(ny):
            J = gj * ny + j
            for i in range(nx):
                I = gi * nx + i
                out.append(_row(k + 1, [xc[I], yc[J], 0.0, 0.0, *(m[J, I] for m in maxima)]))
            out.append("")
        out.append("")
    return out


def write_maxima_txt(path: str, cfg, tiles, maxima) -> None:
    """Write the maxima file at path (replaced if present)."""
    with open(path, "w") as f:
        f.write("\n".join(maxima_lines(cfg, tiles, maxima)) + "\n")
