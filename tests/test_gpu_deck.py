"""The reference's own test decks (tests/golden/decks, copied from its tests/)
through the device path: against the oracle, and the reference's
tile-independence checks (tests/runall.jl:45-46, tests/testlib.jl:363-407):
the same flow on tiles of different sizes gives the same fields."""
import os

import numpy as np
import pytest

from kestrel_amd import deck
from tests.helpers import FIELDS, active_mask, oracle_for, rel_inf, stepper_for

pytestmark = pytest.mark.gpu

DECKS = os.path.join(os.path.dirname(__file__), "golden", "decks")


def load(name):
    return deck.load(os.path.join(DECKS, name))


@pytest.mark.parametrize("name,steps,tol", [
    ("Input_flux_hydro_2d.txt", 60, 0.0),             # Chezy, morphodynamics off: bit for bit
    ("Input_flux_single_pt.txt", 60, 0.0),
    ("Input_cap_morpho_2d.txt", 40, 1e-12),            # Variable drag (tanh switch), Mixed erosion
    ("Input_tile_indep_dynamic_20m.txt", 40, 1e-12),   # tiles of 10 cells, eddy viscosity
])
def test_reference_deck_against_oracle(gpu, name, steps, tol):
    run = load(name)
    o = oracle_for(run.workload)
    g = stepper_for(run.workload)
    so, sg = o.step_n(steps), g.step_n(steps)
    assert sg["rejected"] == so["rejected"] and sg["active_tiles"] == so["active_tiles"]
    a, b = g.state(), o.state()
    assert np.array_equal(a["tile_state"], b["tile_state"])
    assert abs(a["t"] - b["t"]) <= tol * b["t"]
    m = active_mask(run.workload.cfg, b["tile_state"])
    for f in FIELDS:
        r = rel_inf(a[f], b[f], m)
        assert r <= tol, f"{name}: {f} rel {r:.3e}"


def _common(small, big):
    """Slices of the larger grid that hold the smaller, centred grid."""
    cs, cb = small.workload.cfg, big.workload.cfg
    assert cs.dx == cb.dx
    ox, oy = (cb.NX - cs.NX) // 2, (cb.NY - cs.NY) // 2
    return slice(oy, oy + cs.NY), slice(ox, ox + cs.NX)


@pytest.mark.parametrize("small,big,tol", [
    ("Input_tile_indep_static_50m.txt", "Input_tile_indep_static_100m.txt", 1e-13),
    ("Input_tile_indep_dynamic_20m.txt", "Input_tile_indep_dynamic_100m.txt", 1e-11),
])
def test_tile_independence_decks(gpu, small, big, tol):
    """Run both decks to their T end; on the small grid's cells the depth,
    momenta, solids and bed agree to the reference's tolerance."""
    rs, rb = load(small), load(big)
    out = []
    for r in (rs, rb):
        g = stepper_for(r.workload)
        g.integrate_to(r.tend)
        out.append(g.state())
        g.close()
    a, b = out
    sl = _common(rs, rb)
    assert a["t"] == b["t"] == rs.tend
    worst = 0.0
    for f in ("Hn", "rhoHnu", "rhoHnv", "Hnpsi"):
        d = float(np.abs(a[f] - b[f][sl]).max())
        worst = max(worst, d)
        assert d <= tol * max(1.0, float(np.abs(a[f]).max())), f"{f}: {d:.3e}"
    dbt = float(np.abs(a["btv"] - b["btv"][sl[0].start:sl[0].stop + 1, sl[1].start:sl[1].stop + 1]).max())
    assert dbt <= tol, f"bt: {dbt:.3e}"
    assert np.abs(a["Hn"]).max() > 0.01
    print(f"{small} vs {big}: worst field difference {worst:.3e}, bt {dbt:.3e}")
