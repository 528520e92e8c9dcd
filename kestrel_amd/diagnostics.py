"""Diagnostics on a downloaded state: the Volume.txt quantities.

CalculateVolume (src/Output.f90:617-735) Kahan-sums, over the active tiles in
list order and their cells row by row, the flow volume (w - b0 - bt) gam^2,
the mass rho Hn gam, the bed change bt and the solids Hnpsi gam, each times
dx dy.  This is the reference's conservation observable.
"""
from __future__ import annotations

import numpy as np

from .config import KstConfig
from . import scenario as S


def _kahan_add(x, s, c):
    y = x - c
    t = s + y
    c = (t - s) - y
    return t, c


def volume(cfg: KstConfig, b0v: np.ndarray, state: dict) -> dict:
    b0c, btc, bx, by = S.cell_topography(cfg, b0v)
    gam = S.gamma(cfg, bx, by)
    w, hnpsi = state["w"], state["Hnpsi"]
    # derived Hn and rho as ComputeDesingularisedVariables leaves them
    hn = (-btc + (w - b0c)) * gam
    hn = np.where(hn < 0.0, 0.0, hn)
    hp = np.where(hnpsi < 0.0, 0.0, hnpsi)
    eps = cfg.height_threshold
    den = hn * hn + np.maximum(hn * hn, eps * eps)
    psi = np.minimum(2.0 * hn * hp / den, cfg.max_pack)
    rho = cfg.rhow + (cfg.rhos - cfg.rhow) * psi
    vol = mass = bed = solids = 0.0
    cv = cm = cb = cs = 0.0
    ts = np.asarray(state["tile_state"]).reshape(-1)
    nx, ny = cfg.nx_per_tile, cfg.ny_per_tile
    for k in np.nonzero(ts == 1)[0]:
        ti, tj = k % cfg.nx_tiles, k // cfg.nx_tiles
        sl = (slice(tj * ny, (tj + 1) * ny), slice(ti * nx, (ti + 1) * nx))
        dv = ((w[sl] - b0c[sl] - btc[sl]) * gam[sl] * gam[sl]).ravel()
        dm = (rho[sl] * hn[sl] * gam[sl]).ravel()
        db = btc[sl].ravel()
        ds = (hnpsi[sl] * gam[sl]).ravel()
        for q in range(dv.size):
            vol, cv = _kahan_add(dv[q], vol, cv)
            mass, cm = _kahan_add(dm[q], mass, cm)
            bed, cb = _kahan_add(db[q], bed, cb)
            solids, cs = _kahan_add(ds[q], solids, cs)
    a = cfg.dx * cfg.dy
    return {"volume": vol * cfg.dx * cfg.dy, "mass": mass * cfg.dx * cfg.dy,
            "bed": bed * cfg.dx * cfg.dy, "solids": solids * cfg.dx * cfg.dy, "cell_area": a}
