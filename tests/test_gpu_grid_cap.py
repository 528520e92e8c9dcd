"""Launch grids cover a cap on the active set (by default twice the set the
library last read back); when the active set outgrows the cap, workgroups
take further work items in turn.  With the cap fixed far below the active
set (kst_set_grid_cap) every launch takes several passes, through the
XCD-chunked map of every shape; the results must be the bits of the default
grids, for the 16 x 16 and the 16 x 4 sub-block shapes, with and without
morphodynamics."""
import numpy as np
import pytest

from kestrel_amd import workloads
from tests.helpers import FIELDS, stepper_for

pytestmark = pytest.mark.gpu

STATS = ("steps", "rejected", "rhs_evals", "cell_updates", "active_tiles", "t")


def _run(wl, steps, cap):
    g = stepper_for(wl)
    if cap is not None:
        g.set_grid_cap(cap)
    st = g.step_n(steps)
    out = g.state()
    g.close()
    return st, out


@pytest.mark.parametrize("case", ["plain16x16", "plain16x4", "morpho16x4", "morpho16x16"])
def test_small_cap_equals_default(gpu, case):
    if case == "plain16x16":
        # 1.2 k active cells per tile row: more than 128 sub-blocks of 16 x 16
        wl, steps, cap = workloads.plain_cap(512, 0, tile=32, cap_x=0.0, cap_radius=120.0), 40, 2
    elif case == "plain16x4":
        wl, steps, cap = workloads.plain_cap(256, 0, tile=16, cap_radius=20.0), 60, 1
    elif case == "morpho16x4":
        wl, steps, cap = workloads.morpho_cap(128, 0, tile=16, radius=12.0), 40, 1
    else:
        wl, steps, cap = workloads.morpho_cap(512, 0, tile=32, radius=150.0), 12, 2
    st0, s0 = _run(wl, steps, None)
    st1, s1 = _run(wl, steps, cap)
    assert st0["active_tiles"] > 2 * cap
    for k in STATS:
        assert st0[k] == st1[k], f"{case}: {k} {st0[k]} vs {st1[k]}"
    assert np.array_equal(s0["tile_state"], s1["tile_state"])
    for f in FIELDS:
        assert np.array_equal(s0[f], s1[f]), f"{case}: {f} differs with the grid cap at {cap} tiles"


def test_grid_cap_rejects_negative(gpu):
    g = stepper_for(workloads.plain_cap(256, 0, tile=16))
    with pytest.raises(Exception):
        g.set_grid_cap(-1)
    g.close()


@pytest.mark.parametrize("case", ["plain_large", "plain_small", "morpho"])
def test_autotune_same_bits(gpu, case):
    """The autotuner's timed trials (stage-1 launches between saved and
    restored control blocks) and whichever launch shape it picks per
    active-set bucket leave the results bit for bit as the fixed rule's."""
    if case == "plain_large":
        # ~280 active tiles of 32 x 32: the default grid exceeds the resident one
        wl, steps = workloads.plain_cap(1024, 0, tile=32, cap_x=0.0, cap_radius=300.0), 20
    elif case == "plain_small":
        wl, steps = workloads.plain_cap(256, 0, tile=16, cap_radius=20.0), 40
    else:
        wl, steps = workloads.morpho_cap(128, 0, tile=16, radius=12.0), 30
    st0, s0 = _run(wl, steps, None)
    g = stepper_for(wl)
    g.autotune(True)
    st1 = g.step_n(steps)
    s1 = g.state()
    ch = g.autotune_choices()
    g.close()
    assert ch and all(c["stage1_us"] for c in ch), ch
    print(f"{case}: {ch}")
    for k in STATS:
        assert st0[k] == st1[k], f"{case}: {k} {st0[k]} vs {st1[k]}"
    for f in FIELDS:
        assert np.array_equal(s0[f], s1[f]), f"{case}: {f} differs with the autotuner on"
