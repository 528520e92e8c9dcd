"""Activation on the device walks the candidate sides from LDS, or the active
list in place (as the reference's CheckIfNearBoundaries does,
src/TimeStepper.f90:935-1150) when a step has more candidates than
kst_set_activation_candidates allows.  Lowering that cap to one candidate
sends every step with several candidates through the in-place walk; both
runs must agree bit for bit, in the state and the tile states."""
import numpy as np
import pytest

from kestrel_amd import workloads
from tests.helpers import FIELDS, stepper_for

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["plain", "morpho"])
def test_in_place_walk_equals_candidate_walk(gpu, case):
    if case == "plain":
        wl = workloads.plain_cap(256, 0, tile=16)          # 12 -> 14 active tiles in 120 steps
    else:
        wl = workloads.morpho_cap(128, 0, tile=16, radius=10.0)   # 4 -> 10, two tiles at a time
    a, b = stepper_for(wl), stepper_for(wl)
    b.set_activation_candidates(1)
    n0 = int(np.sum(a.state()["tile_state"] == 1))   # KST_TILE_ACTIVE
    for _ in range(6):
        sa, sb = a.step_n(20), b.step_n(20)
        assert sa["rejected"] == sb["rejected"]
        xa, xb = a.state(), b.state()
        assert xa["t"] == xb["t"]
        assert np.array_equal(xa["tile_state"], xb["tile_state"])
        for f in FIELDS:
            assert np.array_equal(xa[f], xb[f]), f
    assert int(np.sum(xa["tile_state"] == 1)) >= n0 + 2   # the active set grew
    b.set_activation_candidates(0)
    with pytest.raises(Exception):
        b.set_activation_candidates(4096)
    a.close()
    b.close()
