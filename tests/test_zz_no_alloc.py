"""kst_volume_noalloc (no device allocation; partial sums in the library's
fixed per-GPU scratch array) against kst_volume through the C-ABI: the same
bits on two calls in a row, the second on a smaller active set, after calls
of kst_volume that left different partials in the handle's own buffer."""
import numpy as np
import pytest

from kestrel_amd import workloads
from tests.helpers import stepper_for

pytestmark = pytest.mark.gpu


def test_volume_noalloc_equals_volume(gpu):
    cases = [(workloads.plain_cap(512, 0, tile=32, cap_x=0.0, cap_radius=120.0), 20),
             (workloads.morpho_cap(128, 0, tile=16, radius=12.0), 10)]
    sizes = []
    for wl, steps in cases:
        g = stepper_for(wl)
        st = g.step_n(steps)
        sizes.append(st["active_tiles"])
        ref = g.volume()
        got = g.volume_noalloc()
        again = g.volume_noalloc()
        g.close()
        assert np.all(np.isfinite(ref)) and ref[0] > 0.0
        assert np.array_equal(got, ref), (wl.name, got, ref)
        assert np.array_equal(again, ref), (wl.name, again, ref)
    assert sizes[1] < sizes[0], sizes
