"""Output files of the reference, written from device results.

CalculateVolume (src/Output.f90:617-735) appends one line per output time to
Volume.txt: the time as F12.2, then six ES24.15 columns separated by ', ',
under a header of right-justified names (src/Output.f90:664-673, 720-727).
The six values come from the device (`Stepper.volume`, i.e. kst_volume).
"""
from __future__ import annotations

import math
import os

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
