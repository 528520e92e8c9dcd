"""Restart from text output on the device: files written from a device run
are read back (kestrel_amd.restart) into a new handle and into the oracle;
both continue bit for bit.  The maxima resume through kst_set_maxima, the
file's cell topography through kst_set_cell_topography.  (The reference's
own restart, from files it wrote, is a run-vector case: restart_bowl_2d.)"""
import numpy as np
import pytest

from kestrel_amd import output, restart, workloads
from tests.helpers import FIELDS, oracle_for, stepper_for

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["plain", "morpho"])
def test_restart_continues_as_the_oracle(gpu, tmp_path, case):
    if case == "plain":
        wl = workloads.plain_cap(128, 0, tile=16, cap_x=0.0, cap_radius=10.0)
    else:
        wl = workloads.morpho_cap(128, 0, tile=16, radius=10.0, erosion="simple", deposition="simple",
                                  transition="step", damping="rat3")
    g = stepper_for(wl)
    st = g.step_n(30)
    tiles, F = g.output_fields()
    p, pm = str(tmp_path / "000001.txt"), str(tmp_path / "Maximums.txt")
    output.write_solution_txt(p, wl.cfg, tiles, F, wl.b0v, g.tile_bt())
    output.write_maxima_txt(pm, wl.cfg, tiles, g.maxima())
    g.close()

    r = stepper_for(wl)
    rs = restart.restart(r, wl.cfg, wl.b0v, st["t"], p, maxima_path=pm)
    mx = r.maxima()
    for q in range(11):
        assert np.array_equal(mx[q][rs.mask], rs.maxima[q][rs.mask])
    o = oracle_for(wl)
    restart.restart(o, wl.cfg, wl.b0v, st["t"], p, maxima_path=pm)
    sr = r.step_n(20)
    so = o.step_n(20)
    assert sr["rejected"] == so["rejected"]
    a, b = r.state(), o.state()
    assert a["t"] == b["t"]
    assert np.array_equal(a["tile_state"], b["tile_state"])
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), f
    r.close()


def test_set_maxima_round_trip(gpu):
    wl = workloads.plain_cap(128, 0, tile=16, cap_x=0.0, cap_radius=10.0)
    g = stepper_for(wl)
    g.step_n(10)
    m0 = g.maxima()
    m1 = [m * 2.0 + 1.0 for m in m0]
    g.set_maxima(m1[:5] + [None] + m1[6:])
    got = g.maxima()
    for q in range(11):
        assert np.array_equal(got[q], m0[q] if q == 5 else m1[q])
    g.close()
