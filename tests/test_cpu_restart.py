"""Restart from text output (LoadInitialCondition_txt, src/Restart.f90:151-328)
on the CPU: the oracle's state written with the reference's formats and read
back by kestrel_amd.restart -- tiles, conserved fields, vertex bt and maxima
to the files' 11 significant digits -- and a restarted run that stays with
the continuous one."""
import numpy as np
import pytest

from kestrel_amd import output, restart, workloads
from tests.helpers import FIELDS, active_mask, oracle_for, oracle_output_fields

REL = 5e-11     # half a unit in the 11th significant digit


def close(a, b):
    return np.abs(a - b) <= REL * np.abs(b) + 1e-300


@pytest.fixture(params=["plain", "morpho"])
def deck(request):
    if request.param == "plain":
        return workloads.plain_cap(128, 0, tile=16, cap_x=0.0, cap_radius=10.0)
    return workloads.morpho_cap(128, 0, tile=16, radius=10.0, erosion="simple", deposition="simple",
                                transition="step", damping="rat3")


def write_files(tmp_path, o, wl):
    tiles, G = oracle_output_fields(o, wl.cfg)
    p = str(tmp_path / "000001.txt")
    output.write_solution_txt(p, wl.cfg, tiles, G, wl.b0v, o.tile_bt())
    pm = str(tmp_path / "Maximums.txt")
    output.write_maxima_txt(pm, wl.cfg, tiles, o.maxima())
    return p, pm


def test_files_read_back(tmp_path, deck):
    wl = deck
    o = oracle_for(wl)
    o.step_n(30)
    p, pm = write_files(tmp_path, o, wl)
    rs = restart.load_txt(wl.cfg, wl.b0v, p, maxima_path=pm)
    act, _ = o.tile_lists()
    assert rs.tiles == act
    st = o.state()
    m = active_mask(wl.cfg, st["tile_state"])
    assert np.array_equal(rs.init.tile_state == 1, st["tile_state"] == 1)
    for f, got in zip(FIELDS, (rs.init.w, rs.init.rhoHnu, rs.init.rhoHnv, rs.init.Hnpsi)):
        assert close(got[m], st[f][m]).all(), f
    # vertex bt: the owner tile's copy, which kst_get_state reports too
    owner = st["btv"]
    assert close(rs.btv, owner).all()
    for got, want in zip(rs.maxima, o.maxima()):
        assert close(got[m], want[m]).all()


def test_restarted_run_stays_with_the_continuous_one(tmp_path, deck):
    wl = deck
    o = oracle_for(wl)
    s = o.step_n(30)
    t = o.state()["t"]
    p, pm = write_files(tmp_path, o, wl)
    rs = restart.load_txt(wl.cfg, wl.b0v, p, maxima_path=pm)
    r = oracle_for(wl)
    r.set_state(rs.init, t=t, btv=rs.btv)
    o.step_n(20)
    r.step_n(20)
    a, b = r.state(), o.state()
    m = active_mask(wl.cfg, b["tile_state"])
    assert np.array_equal(a["tile_state"], b["tile_state"])
    assert abs(a["t"] - b["t"]) <= 1e-9 * b["t"]
    for f in FIELDS:
        scale = np.abs(b[f][m]).max()
        if scale > 0:
            assert np.abs(a[f][m] - b[f][m]).max() <= 1e-7 * scale, f


def test_restart_rejects_a_truncated_file(tmp_path):
    wl = workloads.plain_cap(128, 0, tile=16, cap_x=0.0, cap_radius=10.0)
    o = oracle_for(wl)
    p, _ = write_files(tmp_path, o, wl)
    lines = open(p).read().splitlines()
    open(p, "w").write("\n".join(lines[:-5]) + "\n")
    with pytest.raises(ValueError):
        restart.read_solution_txt(wl.cfg, p)
