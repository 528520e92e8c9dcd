"""CPU tests: the oracle pinned to the reference, invariants the mathematics
fixes, the host-side set-up, and the C-ABI library's exports.

No test here touches a GPU.
"""
import ctypes
import json
import math
import os
import re

import numpy as np
import pytest

from kestrel_amd import config as C
from kestrel_amd import scenario as S, workloads
from oracle.oracle import volume_from_state
from tests.helpers import FIELDS, active_mask, oracle_for, rel_inf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _load_fixture(case):
    z = np.load(os.path.join(GOLDEN, f"rhs_{case}.npz"), allow_pickle=False)
    return z


def _oracle_on_fixture(case, mutate=None):
    from oracle.oracle import Oracle
    wl = workloads.fixture_case(case)
    if mutate:
        mutate(wl.cfg)
    z = _load_fixture(case)
    o = Oracle(wl.cfg)
    o.set_topography(z["b0v"])
    init = S.InitialState(z["w"], z["rhoHnu"], z["rhoHnv"], z["Hnpsi"], z["tile_state"],
                          np.zeros_like(z["w"]))
    o.set_state(init, t=float(z["t"]))
    e, i, adv = o.rhs(1, float(z["t"]))
    return z, e, i, adv


# ---------------------------------------------------------------------------
# the oracle against the reference's own arithmetic
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("case", workloads.FIXTURE_CASES)
def test_oracle_rhs_matches_reference_vectors(case):
    """CalculateHydraulicRHS of the reference (committed vectors, written by
    tests/golden/make_rhs_vectors.py) == the oracle's RHS, bit for bit."""
    z, e, i, adv = _oracle_on_fixture(case)
    assert np.array_equal(e, z["ddte"]), f"{case}: ddtExplicit max diff {np.abs(e - z['ddte']).max()}"
    assert np.array_equal(i, z["ddti"]), f"{case}: ddtImplicit max diff {np.abs(i - z['ddti']).max()}"
    assert adv == float(z["advised"])
    assert np.count_nonzero(z["ddte"]) > 100


def test_fixture_configs_match_deck_builder():
    for case in workloads.FIXTURE_CASES:
        z = _load_fixture(case)
        stored = json.loads(str(z["config"]))
        now = workloads.fixture_case(case).cfg.as_dict()
        for k, v in stored.items():
            assert now[k] == v, (case, k)


@pytest.mark.parametrize("mutation", ["theta", "drag", "gravity"])
def test_planted_mistakes_are_caught(mutation):
    """A plausible slip in the oracle (wrong limiter, wrong drag law, wrong g)
    must fail the reference-vector check."""
    def mutate(cfg):
        if mutation == "theta":
            cfg.limiter = C.LIMITERS["minmod1"]
        elif mutation == "drag":
            cfg.drag = C.DRAGS["voellmy"]
        else:
            cfg.g = 9.80665
    z, e, i, adv = _oracle_on_fixture("chezy", mutate)
    assert not (np.array_equal(e, z["ddte"]) and np.array_equal(i, z["ddti"]))


def test_oracle_rhs_matches_live_reference_with_stale_ghosts():
    """With oracle/_ref built, compare along a run with tile activation, where
    ghost-tile w reconstructions are stale (taken when the tile was added)."""
    from oracle import oracle as OR
    if not OR.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    wl = workloads.plain_cap(128, 0, tile=16, cap_x=0.0, cap_radius=10.0)
    o = oracle_for(wl)
    hs = np.zeros(wl.cfg.ntiles, np.int8)
    checked = 0
    for chunk in range(4):
        o.step_n(40)
        st = o.state()
        f13 = o.fields13()
        sx, sy = o.ghost_slopes()
        e, i, adv = o.rhs(1, st["t"])
        re_, ri, radv = OR.ref_rhs(wl.cfg, st["tile_state"], hs, wl.b0v, f13, sx, sy,
                                   recompute_ghost=False, t=st["t"], grid_t=st["t"])
        assert np.array_equal(e, re_) and np.array_equal(i, ri) and adv == radv
        checked += 1
    assert st["tile_state"].tolist().count(1) > 4


# ---------------------------------------------------------------------------
# invariants fixed by the mathematics
# ---------------------------------------------------------------------------

def _bowl_lake(level=3.0, n_tiles=8, tile=16):
    cfg = C.default_config(n_tiles, n_tiles, tile, tile, dx=1.0)
    C.hydraulic_only(cfg)
    cfg.max_dt = 0.05
    C.derive(cfg)
    xv, yv = S.vertex_coords(cfg)
    X, Y = np.meshgrid(xv, yv)
    b0v = 0.004 * (X * X + Y * Y) + 0.3 * np.sin(0.3 * X) * np.cos(0.2 * Y)
    caps = [S.Cap(0.0, 0.0, 45.0, level, shape="level")]
    init = S.initial_state(cfg, b0v, caps)
    wl = workloads.Workload("bowl_lake", cfg, b0v, init, 0, caps, [])
    return wl


def test_lake_at_rest():
    """A level lake over a bumpy bowl is at rest wherever the stencil is
    submerged: the hydraulic RHS vanishes to round-off there (well-balancing,
    Chertock et al. 2015; the reference checks it on a fully submerged deck,
    tests/Input_lake_at_rest_hydro_2d.txt).  At the shoreline the reference's
    own operator is not balanced (its RHS there is O(100)), so those cells are
    excluded, as in the reference's test."""
    wl = _bowl_lake()
    o = oracle_for(wl)
    st = o.state()
    e, i, adv = o.rhs(1, 0.0)
    wet = st["Hn"] > 0.05
    deep = wet.copy()
    for sh in (1, 2, 3):
        for ax in (0, 1):
            deep &= np.roll(wet, sh, ax) & np.roll(wet, -sh, ax)
    assert deep.sum() > 500
    # momentum balance: hydrostatic flux differences vs gravity, each ~1e4
    assert np.abs(e[1][deep]).max() < 1e-9
    assert np.abs(e[2][deep]).max() < 1e-9
    assert np.abs(e[0][deep]).max() < 1e-13
    assert np.abs(i[1][deep]).max() == 0.0


def test_mass_conservation_and_positivity():
    """Volume by the Volume.txt definition is conserved with no sources and a
    closed domain (reference tolerance 1e-10, tests/testlib.jl:113), and the
    depth stays non-negative (tests/testlib.jl:286-302)."""
    wl = workloads.plain_cap(128, 0, tile=16, cap_x=0.0, cap_radius=10.0)
    o = oracle_for(wl)
    v0 = volume_from_state(wl.cfg, wl.b0v, o.state())["volume"]
    for _ in range(3):
        o.step_n(50)
        st = o.state()
        v = volume_from_state(wl.cfg, wl.b0v, st)["volume"]
        assert abs(v - v0) / v0 < 1e-10, (v, v0)
        m = active_mask(wl.cfg, st["tile_state"])
        b0c, btc, bx, by = S.cell_topography(wl.cfg, wl.b0v)
        hp = (st["w"] - b0c)[m]
        assert hp.min() >= -1e-14


def test_symmetry_about_the_x_axis():
    """A cap centred on y = 0 over an x-only bed stays mirror-symmetric."""
    wl = workloads.plain_cap(128, 0, tile=16, cap_x=0.0, cap_radius=10.0)
    o = oracle_for(wl)
    o.step_n(60)
    st = o.state()
    assert rel_inf(st["w"], st["w"][::-1, :]) < 1e-12
    assert rel_inf(st["rhoHnu"], st["rhoHnu"][::-1, :]) < 1e-12
    assert rel_inf(st["rhoHnv"], -st["rhoHnv"][::-1, :]) < 1e-12


def test_tile_independence():
    """Two tile layouts of the same grid agree (tests/runall.jl:45-46: 1e-13
    with a static tile set, 1e-11 with activation)."""
    res = []
    for tile in (16, 32):
        wl = workloads.plain_cap(128, 0, tile=tile, cap_x=0.0, cap_radius=10.0)
        o = oracle_for(wl)
        o.step_n(40)
        res.append(o.state())
    a, b = res
    both = (a["Hn"] > 1e-6) | (b["Hn"] > 1e-6)
    for f in ("Hn", "Hnpsi"):
        assert np.abs(a[f] - b[f])[both].max() <= 1e-11
    ua = np.where(a["Hn"] > 1e-3, a["rhoHnu"] / np.maximum(a["Hn"], 1e-3), 0)
    ub = np.where(b["Hn"] > 1e-3, b["rhoHnu"] / np.maximum(b["Hn"], 1e-3), 0)
    assert np.abs(ua - ub).max() <= 1e-8


def test_limiters_closed_forms():
    """The oracle's limiters (src/Limiters.f90:83-184) against their textbook
    definitions, evaluated independently here over slopes of both signs,
    zeros and a wide range of magnitudes: MinMod (Roe), generalised MinMod
    with theta = 1.3, van Albada, the WENO-weighted average and the central
    difference; plus the properties a mistake would break (symmetry,
    consistency lim(a, a) = a, zero at a sign change for the MinMods, the
    TVD bounds, the weighted average lying between its arguments)."""
    from oracle.oracle import lib
    L = lib()
    rng = np.random.default_rng(7)
    vals = np.concatenate([rng.standard_normal(400) * 10.0 ** rng.integers(-6, 4, 400),
                           [0.0, -0.0, 1.0, -1.0, 2.5, -2.5, 1e-300]])
    A, B = np.meshgrid(vals[::7], vals[::5])
    A, B = A.ravel(), B.ravel()
    lim = {name: np.array([L.ko_limiter(C.LIMITERS[name], a, b) for a, b in zip(A, B)])
           for name in ("minmod1", "minmod2", "vanalbada", "weno", "none")}
    same = A * B > 0
    sgn = np.sign(A)
    # MinMod: the smaller magnitude with the common sign, else 0
    mm1 = np.where(same, sgn * np.minimum(np.abs(A), np.abs(B)), 0.0)
    assert np.array_equal(lim["minmod1"], mm1)
    # generalised MinMod, theta = 1.3: min(theta|a|, theta|b|, |a+b|/2), common sign
    th = 1.3
    mm2 = np.where(same, sgn * np.minimum(np.minimum(np.abs(th * A), np.abs(th * B)),
                                           np.abs(0.5 * (A + B))), 0.0)
    assert np.array_equal(lim["minmod2"], mm2)
    assert np.all(np.abs(lim["minmod2"]) <= th * np.minimum(np.abs(A), np.abs(B)))
    # van Albada: ab(a + b) / (a^2 + b^2), 0 where both vanish
    den = A * A + B * B
    with np.errstate(invalid="ignore", divide="ignore"):
        va = np.where(den == 0.0, 0.0, (A * A * B + A * B * B) / den)
    assert np.allclose(lim["vanalbada"], va, rtol=1e-14, atol=0.0)
    # WENO weights 1/(s^2 + eps)^2: an average that lies between a and b
    ea, eb = A * A + 1e-6, B * B + 1e-6
    wa, wb = 1.0 / (ea * ea), 1.0 / (eb * eb)
    assert np.allclose(lim["weno"], (wa * A + wb * B) / (wa + wb), rtol=1e-14, atol=0.0)
    lo, hi = np.minimum(A, B), np.maximum(A, B)
    eps = 1e-15 * np.maximum(np.abs(lo), np.abs(hi))
    assert np.all((lim["weno"] >= lo - eps) & (lim["weno"] <= hi + eps))
    # no limiter: the central difference
    assert np.array_equal(lim["none"], 0.5 * (A + B))
    # symmetry and consistency
    for name, f in lim.items():
        g = np.array([L.ko_limiter(C.LIMITERS[name], b, a) for a, b in zip(A, B)])
        if name in ("minmod1", "minmod2", "none"):
            assert np.array_equal(f, g), name
        else:   # the two orders round differently: a few ulps of the slopes
            assert np.all(np.abs(f - g) <= 4e-16 * np.maximum(np.abs(A), np.abs(B))), name
        for a in vals[:50]:
            assert L.ko_limiter(C.LIMITERS[name], a, a) == pytest.approx(a, rel=1e-15, abs=0.0), name
    # the MinMods vanish at a sign change
    for name in ("minmod1", "minmod2"):
        assert np.all(lim[name][A * B <= 0] == 0.0), name


# ---------------------------------------------------------------------------
# host-side set-up and the C-ABI
# ---------------------------------------------------------------------------

def test_scenario_caps_match_oracle_painting():
    """numpy cap painting == the oracle's LoadSourceConditions restatement."""
    wl = workloads.plain_cap(256, 0)
    from oracle.oracle import Oracle
    o = Oracle(wl.cfg)
    o.set_topography(wl.b0v)
    o.add_caps(wl.caps)
    o.finish_sources()
    st = o.state()
    assert np.array_equal(st["tile_state"], wl.init.tile_state)
    m = active_mask(wl.cfg, st["tile_state"])
    for f in FIELDS:
        assert np.array_equal(st[f][m], getattr(wl.init, f)[m]), f
    mx = o.maxima()
    assert np.array_equal(mx[0][m], wl.init.hnmax0[m])


def test_flux_source_cell_count_matches_oracle():
    cfg = C.default_config(6, 6, 16, 16, dx=5.0)
    C.hydraulic_only(cfg)
    src = C.FluxSource(2.5, 2.5, 7.0, [0.0, 10.0], [50.0, 20.0], [0.0, 0.1])
    b0v = S.topo_flat(cfg)
    init = S.initial_state(cfg, b0v, [], [src])
    C.set_sources(cfg, [src])
    from oracle.oracle import Oracle
    o = Oracle(cfg)
    o.set_topography(b0v)
    n = o.add_flux_sources()
    assert n == [src.num_cells] and src.num_cells > 1
    assert np.array_equal(o.state()["tile_state"], init.tile_state)


def test_config_defaults_match_library():
    from kestrel_amd import runtime
    L = runtime.lib()
    c = C.KstConfig()
    L.kst_config_defaults(ctypes.byref(c))
    py = C.default_config(4, 4)
    for k, v in c.as_dict().items():
        if k in ("nx_tiles", "ny_tiles", "nx_per_tile", "ny_per_tile", "dx", "dy", "x_size", "y_size",
                 "diffusive_time_scale"):
            continue
        assert v == py.as_dict()[k], k


def test_library_exports_every_declared_symbol():
    """libkestrel_amd.so loads (no device needed) and exports every function
    include/kestrel_step.h declares."""
    from kestrel_amd import runtime
    L = runtime.lib()
    hdr = open(os.path.join(ROOT, "include", "kestrel_step.h")).read()
    names = set(re.findall(r"^\s*(?:int|void)\s+(kst_\w+)\s*\(", hdr, re.M))
    assert len(names) >= 13
    for n in names:
        assert hasattr(L, n), n
        assert n in runtime.EXPORTS, n


def test_product_has_no_oracle_dependency():
    """The package never imports or loads the oracle."""
    pkg = os.path.join(ROOT, "kestrel_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h", ".f90")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.replace("oracle's", ""), os.path.join(dirpath, f)


def test_conditioning_probe_moves_only_transcendental_paths():
    """oracle/liboracle_pert.so (per-cell tanh, cos, pow, log one ulp up)
    gives the plain oracle's bits on a deck that evaluates none of them
    (Chezy drag, no morphodynamics) and differs, slightly, on one that does
    (Variable drag with erosion and deposition)."""
    import numpy as np
    from oracle.oracle import Oracle
    from tests import runvec

    def run(name, perturbed, tend):
        r = runvec.load_case(name)
        o = Oracle(r.workload.cfg, perturbed=perturbed)
        o.set_topography(r.workload.b0v)
        o.set_state(r.workload.init, t=r.tstart)
        o.integrate_to(tend)
        f = o.fields13()
        o.close()
        return f

    a, b = run("lake_at_rest_hydro_2d", False, 0.5), run("lake_at_rest_hydro_2d", True, 0.5)
    assert np.array_equal(a, b)
    a, b = run("flat_depositional_2d", False, 20.0), run("flat_depositional_2d", True, 20.0)
    d = np.abs(a[3] - b[3]).max() / np.abs(a[3]).max() if np.abs(a[3]).max() > 0 else 0.0
    assert not np.array_equal(a, b)
    assert np.abs(a[0] - b[0]).max() < 1e-12 and d < 1e-6
