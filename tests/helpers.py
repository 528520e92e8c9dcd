"""Shared helpers for the parity tests."""
from __future__ import annotations

import numpy as np

FIELDS = ("w", "rhoHnu", "rhoHnv", "Hnpsi")


def rel_inf(a: np.ndarray, b: np.ndarray, mask=None) -> float:
    """max|a-b| / max|b| over the mask (the per-field parity metric)."""
    if mask is not None:
        a = a[mask]
        b = b[mask]
    den = np.max(np.abs(b)) if b.size else 0.0
    num = np.max(np.abs(a - b)) if b.size else 0.0
    if den == 0.0:
        return float(num)
    return float(num / den)


def active_mask(cfg, tile_state) -> np.ndarray:
    ts = np.asarray(tile_state).reshape(cfg.ny_tiles, cfg.nx_tiles)
    return np.repeat(np.repeat(ts == 1, cfg.ny_per_tile, 0), cfg.nx_per_tile, 1)


def bitwise_equal(a: np.ndarray, b: np.ndarray) -> bool:
    """Equal as values (+0 == -0), NaN-free."""
    return bool(np.array_equal(a, b))


def oracle_for(wl):
    from oracle.oracle import Oracle
    o = Oracle(wl.cfg)
    o.set_topography(wl.b0v)
    o.set_state(wl.init)
    return o


def stepper_for(wl):
    from kestrel_amd.runtime import Stepper
    s = Stepper(wl.cfg)
    s.set_topography(wl.b0v)
    s.set_state(wl.init)
    return s


# solution-file column (6-19, 0-based here) -> oracle fields13 index
# (W RHU RHV HNPSI HN UU VV PSI RHO B0 BT DBDX DBDY)
OUT_TO_F13 = {0: 4, 2: 5, 3: 6, 4: 8, 5: 7, 6: 1, 7: 2, 8: 3, 9: 0, 10: 9, 11: 10, 12: 11, 13: 12}


def oracle_output_fields(o, cfg):
    """(tiles, fields[n, 14, ny, nx]) of the solution file from the oracle's
    tileContainer (speed: FlowSquaredSpeedSlopeAligned, src/Closures.f90:178-205)."""
    act, _ = o.tile_lists()
    f13 = o.fields13()
    nx, ny, ntx = cfg.nx_per_tile, cfg.ny_per_tile, cfg.nx_tiles
    G = np.zeros((len(act), 14, ny, nx))
    for q, k in enumerate(act):
        gi, gj = k % ntx, k // ntx
        sl = (slice(gj * ny, (gj + 1) * ny), slice(gi * nx, (gi + 1) * nx))
        for col, idx in OUT_TO_F13.items():
            G[q, col] = f13[idx][sl]
        u, v, bx, by, hn = G[q, 2], G[q, 3], G[q, 12], G[q, 13], G[q, 0]
        m = u * u * (1.0 + bx * bx)
        m = m + v * v * (1.0 + by * by) + 2.0 * bx * by * u * v
        with np.errstate(invalid="ignore"):
            G[q, 1] = np.where(hn > cfg.height_threshold, np.sqrt(m), 0.0)
    return act, G
