"""kst_volume -- CalculateVolume (src/Output.f90:617-735) on the device --
against the oracle's sequential Kahan sums, against the exact sums of the
per-cell terms (math.fsum), against a closed form, and through the
reference's own conservation check (tests/testlib.jl:104-158) at full size.

The device sums are compensated (Sum2), the reference's are Kahan sums: both
lie within an ulp or two of the exact sum, so the bar is the exact sum to one
ulp for the device and three ulps between the two."""
import math

import numpy as np
import pytest

from kestrel_amd import config as C, output, workloads
from tests.helpers import active_mask, oracle_for, stepper_for
from tests.test_cpu_output import flat_lake

pytestmark = pytest.mark.gpu


def ulps(a, b, ref):
    return abs(a - b) / math.ulp(abs(ref)) if ref != 0.0 else abs(a - b)


def device_terms(g, cfg, st):
    """The four per-cell terms from the device's own state and cell
    topography, in numpy (each term is an elementwise IEEE expression)."""
    topo = g.cell_topography()
    b0c, btc, bx, by = topo["b0c"], topo["btc"], topo["dbdx"], topo["dbdy"]
    gam = np.sqrt(1.0 + bx * bx + by * by) if cfg.geometric_factors else np.ones_like(bx)
    w, hp = st["w"], st["Hnpsi"]
    hn = np.maximum((-btc + (w - b0c)) * gam, 0.0)
    hpc = np.maximum(hp, 0.0)
    eps = cfg.height_threshold
    psi = np.minimum(2.0 * hn * hpc / (hn * hn + np.maximum(hn * hn, eps * eps)), cfg.max_pack)
    rho = cfg.rhow + (cfg.rhos - cfg.rhow) * psi
    m = active_mask(cfg, st["tile_state"])
    return [((w - b0c - btc) * gam * gam)[m], (rho * hn * gam)[m], btc[m], (hp * gam)[m]]


@pytest.mark.parametrize("case", ["plain", "morpho", "eddy"])
def test_volume_against_oracle(gpu, case):
    if case == "plain":
        wl, steps = workloads.plain_cap(256, 0, tile=32, cap_x=0.0, cap_radius=30.0), 120
    elif case == "morpho":
        # closures without transcendentals, so device and oracle states agree bitwise
        wl, steps = workloads.morpho_cap(128, 0, tile=16, radius=10.0, erosion="simple",
                                         deposition="simple", transition="step", damping="rat3"), 60
    else:
        wl, steps = workloads.plain_cap(128, 0, tile=16, cap_x=0.0, cap_radius=10.0), 60
        wl.cfg.eddy_viscosity = 0.01
        C.derive(wl.cfg)
    g = stepper_for(wl)
    o = oracle_for(wl)
    for _ in range(2):
        g.step_n(steps)
        o.step_n(steps)
        for f in ("w", "Hnpsi"):
            assert np.array_equal(g.state()[f], o.state()[f])
        dev = g.volume()
        ora = o.volume()
        terms = o.volume_terms()
        a = wl.cfg.dx * wl.cfg.dy
        assert a == 1.0
        for col, q, scale in ((0, 0, 1.0), (2, 1, 1.0), (4, 3, wl.cfg.rhos)):
            exact = math.fsum(terms[:, q])
            assert ulps(dev[col] / scale, exact, exact) <= 1, (case, col, dev[col], exact)
            assert ulps(dev[col], ora[col], ora[col]) <= 3, (case, col, dev[col], ora[col])
        bed_abs = float(np.abs(terms[:, 2]).sum())
        assert abs(dev[1] - math.fsum(terms[:, 2])) <= 2 * math.ulp(bed_abs) + 0.0
        for col in (3, 5):
            assert abs(dev[col] - ora[col]) <= 4 * math.ulp(max(abs(ora[col]), bed_abs * 3000.0))
        # same bits on a repeated call
        assert np.array_equal(dev, g.volume())
    g.close()


def test_flat_lake_closed_form(gpu):
    wl = flat_lake()
    g = stepper_for(wl)
    st = g.state()
    n_wet = int((st["w"] > 0.0).sum())
    v = g.volume()
    assert v[0] == 0.5 * n_wet
    assert v[2] == wl.cfg.rhow * 0.5 * n_wet
    assert v[1] == 0.0 and v[4] == 0.0
    g.close()


def test_mixed_precision_handle_reads_the_current_state(gpu):
    """The mixed path keeps u^n in one of two buffers by step parity."""
    wl = workloads.plain_cap(256, 0, tile=32, cap_x=0.0, cap_radius=30.0)
    wl.cfg.precision = C.PREC_MIXED
    g = stepper_for(wl)
    for n in (1, 1, 7, 10):
        g.step_n(n)
        st = g.state()
        t = device_terms(g, wl.cfg, st)
        v = g.volume()
        for col, q in ((0, 0), (2, 1)):
            exact = math.fsum(t[q])
            assert ulps(v[col], exact, exact) <= 1, (n, col)
    g.close()


def test_conservation_at_full_size(gpu, tmp_path):
    """The reference's check (flow + bed volume, solids mass, 1e-10 relative,
    no flux sources) on device volumes written to and read back from
    Volume.txt: the plain 2048x2048 cap (1.6 M active cells) over 200 steps
    and a 512x512 morphodynamic cap over 150 steps.  Then the 1024x1024
    morphodynamic hills (flux-fed, so not conserved) over 300 steps: the
    device sums within a few ulps of the exact sums of the device state's
    terms, as in every case."""
    cases = ((workloads.plain_cap_2048(200), 200, True),
             (workloads.morpho_cap(512, 0, tile=32, radius=150.0), 150, True),
             (workloads.morpho_hills_1024(300), 300, False))
    for wl, steps, conserved in cases:
        g = stepper_for(wl)
        p = str(tmp_path / f"{wl.name}_Volume.txt")
        v0 = g.volume()
        output.write_volume(p, 0.0, v0, create=True)
        st = g.step_n(steps)
        v1 = g.volume()
        output.write_volume(p, st["t"], v1)
        lines = open(p).read().splitlines()
        first = [float(x) for x in lines[1].split(",")]
        last = [float(x) for x in lines[-1].split(",")]
        # ES24.15 keeps 16 significant digits: half a unit in the 16th
        for got, want in ((first[1:], v0), (last[1:], v1)):
            for a_, b_ in zip(got, want):
                assert abs(a_ - b_) <= 5e-16 * abs(b_) + 1e-300, (a_, b_)
        if conserved:
            vol_err = abs((first[1] + first[2]) - (last[1] + last[2])) / (first[1] + first[2])
            assert vol_err < 1e-10, (wl.name, vol_err)
            if first[5] + first[6] > 0.0:
                sol_err = abs((first[5] + first[6]) - (last[5] + last[6])) / (first[5] + first[6])
                assert sol_err < 1e-10, (wl.name, sol_err)
            print(wl.name, "volume", v1[0], "drift", vol_err)
        t = device_terms(g, wl.cfg, g.state())
        a = wl.cfg.dx * wl.cfg.dy
        for col, q in ((0, 0), (2, 1)):
            exact = math.fsum(t[q]) * a
            assert ulps(v1[col], exact, exact) <= 3, (wl.name, col, v1[col], exact)
        g.close()


# kst_get_output column -> oracle fields13 index (W RHU RHV HNPSI HN UU VV PSI RHO B0 BT DBDX DBDY)
OUT_TO_F13 = {0: 4, 2: 5, 3: 6, 4: 8, 5: 7, 6: 1, 7: 2, 8: 3, 9: 0, 10: 9, 11: 10, 12: 11, 13: 12}


@pytest.mark.parametrize("case", ["set_state", "plain", "morpho", "tile_size"])
def test_output_fields_against_oracle(gpu, case, tmp_path):
    """OutputSolutionData_txt's columns from the device against tileContainer's
    u in the oracle, bit for bit -- velocities included, which the reference
    writes as its last RHS derived them, before the final implicit drag
    update -- and the files written from both the same text."""
    if case == "morpho":
        wl, steps = workloads.morpho_cap(128, 0, tile=16, radius=10.0, erosion="simple",
                                         deposition="simple", transition="step", damping="rat3"), 40
    elif case == "tile_size":
        wl, steps = workloads.plain_cap(120, 0, tile=8, cap_x=0.0, cap_radius=10.0), 60
    else:
        wl, steps = workloads.plain_cap(128, 0, tile=16, cap_x=0.0, cap_radius=10.0), \
            (0 if case == "set_state" else 50)
    g = stepper_for(wl)
    o = oracle_for(wl)
    if steps:
        g.step_n(steps)
        o.step_n(steps)
    tiles, F = g.output_fields()
    act, _ = o.tile_lists()
    assert list(tiles) == act
    f13 = o.fields13()
    cfg = wl.cfg
    nx, ny, ntx = cfg.nx_per_tile, cfg.ny_per_tile, cfg.nx_tiles
    G = np.zeros_like(F)
    for q, k in enumerate(act):
        gi, gj = k % ntx, k // ntx
        sl = (slice(gj * ny, (gj + 1) * ny), slice(gi * nx, (gi + 1) * nx))
        for col, idx in OUT_TO_F13.items():
            G[q, col] = f13[idx][sl]
        # FlowSquaredSpeedSlopeAligned (src/Closures.f90:178-205) of the stored u, v
        u, v, bx, by, hn = G[q, 2], G[q, 3], G[q, 12], G[q, 13], G[q, 0]
        m = u * u * (1.0 + bx * bx)
        m = m + v * v * (1.0 + by * by) + 2.0 * bx * by * u * v
        with np.errstate(invalid="ignore"):
            G[q, 1] = np.where(hn > cfg.height_threshold, np.sqrt(m), 0.0)
    for col in range(14):
        assert np.array_equal(F[:, col], G[:, col]), (case, col)
    if case in ("plain", "tile_size"):
        # the velocities are not the final momentum's: the drag update came after
        hn, ru, rho = F[:, 0], F[:, 6], F[:, 4]
        eps = cfg.height_threshold
        fresh = 2.0 * hn * ru / (hn * hn + np.maximum(hn * hn, eps * eps)) / rho
        assert np.any(fresh != F[:, 2])
    pd, po = str(tmp_path / "dev.txt"), str(tmp_path / "ora.txt")
    output.write_solution_txt(pd, cfg, tiles, F, wl.b0v, g.tile_bt())
    output.write_solution_txt(po, cfg, act, G, wl.b0v, o.tile_bt())
    assert open(pd).read() == open(po).read()
    assert open(pd + "_topo").read() == open(po + "_topo").read()
    g.close()


def test_output_rejects_a_short_buffer(gpu):
    wl = workloads.plain_cap(128, 0, tile=16, cap_x=0.0, cap_radius=10.0)
    g = stepper_for(wl)
    tiles, _ = g.output_fields()
    with pytest.raises(Exception):
        g.output_fields(max_tiles=len(tiles) - 1)
    g.close()


@pytest.mark.parametrize("case", ["plain", "morpho"])
def test_maxima_file_matches_oracle(gpu, tmp_path, case):
    """OutputMaximums_txt from the device maxima and the oracle's: the same text."""
    if case == "morpho":
        wl = workloads.morpho_cap(128, 0, tile=16, radius=10.0, erosion="simple",
                                  deposition="simple", transition="step", damping="rat3")
    else:
        wl = workloads.plain_cap(128, 0, tile=16, cap_x=0.0, cap_radius=10.0)
    g = stepper_for(wl)
    o = oracle_for(wl)
    g.step_n(40)
    o.step_n(40)
    tiles, _ = g.output_fields()
    act, _ = o.tile_lists()
    pd, po = str(tmp_path / "dev_max.txt"), str(tmp_path / "ora_max.txt")
    output.write_maxima_txt(pd, wl.cfg, tiles, g.maxima())
    output.write_maxima_txt(po, wl.cfg, act, o.maxima())
    text = open(pd).read()
    assert text == open(po).read()
    assert len(text.splitlines()) == 1 + len(act) * (wl.cfg.ny_per_tile * (wl.cfg.nx_per_tile + 1) + 1)
    g.close()
