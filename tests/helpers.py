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
