"""Mixed precision (KST_PREC_MIXED) against the FP64 oracle and the FP64
device path.

The mixed path keeps u^n, the w slopes, CorrectSlopes, the face depth, the
hydrostatic flux, the sources and every RK accumulation in FP64, and stores
the RHS and the stage increments over u^n as float, with the Hnpsi, u, v
reconstruction, the wave speeds and the other fluxes in float.  Bar: relative
L2 error of the depth h (and of each conserved field) at most 1e-5 against
FP64 over the active cells, a lake at rest kept at rest to round-off, mass
drift below 1e-10 (the float increments do not telescope exactly; FP64
drifts below 1e-12).
"""
import math

import numpy as np
import pytest

from kestrel_amd import config as C
from kestrel_amd import scenario as S
from kestrel_amd import workloads
from tests.helpers import FIELDS, active_mask, oracle_for, stepper_for

pytestmark = pytest.mark.gpu

TOL_L2 = 1e-5


def mixed(wl):
    wl.cfg.precision = C.PREC_MIXED
    return wl


def rel_l2(a, b, m):
    d = np.linalg.norm((a - b)[m])
    n = np.linalg.norm(b[m])
    return float(d / n) if n > 0 else float(d)


def depth(wl, st):
    b0c, btc, bx, by = S.cell_topography(wl.cfg, wl.b0v)
    return np.maximum(st["w"] - b0c - btc, 0.0)


def errors(wl, a, b):
    m = active_mask(wl.cfg, b["tile_state"])
    out = {f: rel_l2(a[f], b[f], m) for f in FIELDS}
    out["h"] = rel_l2(depth(wl, a), depth(wl, b), m)
    out["h_max"] = float(np.abs(depth(wl, a) - depth(wl, b))[m].max() / depth(wl, b)[m].max())
    return out


def volume(wl, st):
    """Flow volume by the Volume.txt definition ((w - b0 - bt) gam^2 dx dy over
    active cells), summed exactly."""
    b0c, btc, bx, by = S.cell_topography(wl.cfg, wl.b0v)
    gam = S.gamma(wl.cfg, bx, by)
    m = active_mask(wl.cfg, st["tile_state"])
    return math.fsum(((st["w"] - b0c - btc) * gam * gam)[m].ravel()) * wl.cfg.dx * wl.cfg.dy


def test_mixed_against_oracle(gpu):
    """256x256 cap on the 10 degree plane, 300 steps: mixed device path
    against the FP64 oracle."""
    wl = workloads.plain_cap(256, 300)
    o = oracle_for(wl)
    g = stepper_for(mixed(workloads.plain_cap(256, 300)))
    so = o.step_n(300)
    sg = g.step_n(300)
    assert sg["steps"] == so["steps"] == 300
    a, b = g.state(), o.state()
    assert np.array_equal(a["tile_state"], b["tile_state"])
    err = errors(wl, a, b)
    print("mixed vs oracle, 300 steps:", {k: f"{v:.2e}" for k, v in err.items()})
    for k in (*FIELDS, "h"):
        assert err[k] <= TOL_L2, (k, err[k])
    assert abs(sg["t"] - so["t"]) <= 1e-6 * so["t"]


def test_mixed_precision_study_2048(gpu):
    """The precision study at 2048x2048 (1.6 M active cells), 200 steps:
    per-field error of the mixed path against the FP64 device path (itself
    bit-identical to the oracle), and mass drift of both."""
    wf = workloads.plain_cap_2048(200)
    wm = mixed(workloads.plain_cap_2048(200))
    f = stepper_for(wf)
    mx = stepper_for(wm)
    v0 = volume(wf, f.state())
    f.step_n(200)
    mx.step_n(200)
    a, b = mx.state(), f.state()
    err = errors(wf, a, b)
    dv_f = abs(volume(wf, b) - v0) / v0
    dv_m = abs(volume(wf, a) - v0) / v0
    print("mixed vs FP64 at 2048^2, 200 steps:", {k: f"{v:.2e}" for k, v in err.items()},
          f"mass drift fp64 {dv_f:.2e} mixed {dv_m:.2e}")
    for k in (*FIELDS, "h"):
        assert err[k] <= TOL_L2, (k, err[k])
    assert dv_f < 1e-12
    assert dv_m < 1e-10


def lake_on_plane(n=2048):
    """A level lake (w = 100 m) on the study's 10 degree plane, painted in a
    disc of radius 0.3 L at the domain centre (the plane's own shoreline is at
    x = 567 m; beyond the disc the bed below 100 m is dry, so the disc edge
    is a dam, far from the cells measured)."""
    wl = workloads.plain_cap(n, 1, cap_x=0.0, cap_radius=0.3 * n)
    caps = [S.Cap(0.0, 0.0, 0.3 * n, 100.0, shape="level")]
    wl.init = S.initial_state(wl.cfg, wl.b0v, caps)
    wl.caps = caps
    return wl


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_lake_at_rest_on_plane(gpu, prec):
    """One step of the lake: wherever the stencils of the step stay inside the
    lake (at least 16 cells from its shore and from the disc edge) w stays
    100 m and the momentum zero to round-off."""
    wl = lake_on_plane()
    if prec == "mixed":
        mixed(wl)
    g = stepper_for(wl)
    g.step_n(1)
    st = g.state()
    b0c, btc, bx, by = S.cell_topography(wl.cfg, wl.b0v)
    x, y = S.cell_centres(wl.cfg)
    X, Y = np.meshgrid(x, y)
    R = np.hypot(X, Y)
    deep = (R < 0.3 * 2048 - 16 * wl.cfg.dx) & (100.0 - b0c > 16 * wl.cfg.dx * math.tan(math.radians(10)))
    deep &= active_mask(wl.cfg, st["tile_state"])
    assert deep.sum() > 100000
    dw = np.abs(st["w"] - 100.0)[deep].max()
    mom = max(np.abs(st["rhoHnu"])[deep].max(), np.abs(st["rhoHnv"])[deep].max())
    print(f"lake {prec}: max |w - 100| {dw:.2e} m, max |rho Hn u| {mom:.2e}")
    assert dw < 1e-11
    assert mom < 1e-8


def test_mixed_repeatable(gpu):
    """Two mixed runs of the same deck agree bit for bit."""
    runs = []
    for _ in range(2):
        g = stepper_for(mixed(workloads.plain_cap(256, 50)))
        g.step_n(50)
        runs.append(g.state())
    for f in FIELDS:
        assert np.array_equal(runs[0][f], runs[1][f]), f
