"""Input decks: the reference's own test decks (tests/golden/decks, copied
from the reference's tests/) read into this library's configuration as the
reference's setters would (src/Input.f90, src/Parameters.f90,
src/SolverSettings.f90, src/DomainSettings.f90, src/InitConds.f90,
src/TopogFuncs.f90)."""
import glob
import math
import os

import numpy as np
import pytest

from kestrel_amd import config as C
from kestrel_amd import deck
from kestrel_amd import scenario as S

DECKS = os.path.join(os.path.dirname(__file__), "golden", "decks")

# every reference deck loads (the 9 one-dimensional ones take the 1D branches)
EXPECT = {}


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(DECKS, "Input_*.txt"))),
                         ids=lambda p: os.path.basename(p))
def test_reference_decks_load_or_name_what_is_unsupported(path):
    name = os.path.basename(path)
    if name in EXPECT:
        with pytest.raises(deck.DeckUnsupported, match=EXPECT[name]):
            deck.load(path)
        return
    run = deck.load(path)
    c = run.workload.cfg
    assert c.NX == c.nx_tiles * c.nx_per_tile and run.tend > 0
    assert (run.workload.init.tile_state == 1).any()
    if c.ny_tiles * c.ny_per_tile == 1:
        assert run.workload.init.w.shape == (1, c.NX) and run.workload.b0v.shape == (2, c.NX + 1)


def test_cap_morpho_deck_values():
    run = deck.load(os.path.join(DECKS, "Input_cap_morpho_2d.txt"))
    c = run.workload.cfg
    assert (c.nx_tiles, c.ny_tiles, c.nx_per_tile, c.ny_per_tile) == (40, 40, 50, 50)
    assert c.dx == c.dy == 1.0 and c.x_size == 2000.0
    assert c.drag == C.DRAGS["variable"] and c.chezy_co == 0.04
    assert (c.pouliquen_min, c.pouliquen_max, c.pouliquen_beta) == (0.1, 0.4, 0.126)
    assert c.erosion == C.EROSIONS["mixed"] and c.morphodynamics == 1
    assert (c.ero_rate, c.ero_rate_granular, c.ero_depth) == (1e-3, 0.1, 5.0)
    assert (c.voellmy_switch_rate, c.voellmy_switch_value) == (3.0, 0.2)
    assert (c.bed_porosity, c.max_pack, c.rhow, c.rhos) == (0.35, 0.65, 1000.0, 2000.0)
    assert (c.ero_critical_height, c.solid_diameter) == (0.01, 1e-3)
    assert (c.cfl, c.max_dt, c.tile_buffer, c.height_threshold) == (0.25, 10.0, 3, 1e-5)
    assert run.tend == 5.0 and run.nout == 5
    # defaults the deck leaves out (src/Parameters.f90:41-77)
    assert c.deposition == C.DEPOSITIONS["spearman manning"]
    assert c.ero_transition == C.TRANSITIONS["smooth"] and c.fswitch == C.SWITCHES["tanh"]
    assert c.geometric_factors == 1 and c.eddy_viscosity == 0.0 and c.limiter == C.LIMITERS["minmod2"]
    cap = run.workload.caps[0]
    assert (cap.x, cap.y, cap.radius, cap.height, cap.psi, cap.shape) == (0.0, 100.0, 10.0, 1.0, 0.1, "flat")
    xv, yv = S.vertex_coords(c)
    assert np.array_equal(run.workload.b0v, np.broadcast_to(-0.2 * xv[None, :], run.workload.b0v.shape))
    # the deck names 'Pouliquen L', which the reference does not read either
    assert any("pouliquen l" in w for w in run.deck.warnings)


def test_flux_source_deck_values():
    run = deck.load(os.path.join(DECKS, "Input_flux_single_pt.txt"))
    c = run.workload.cfg
    assert c.morphodynamics == 0 and c.n_sources == 1
    s = run.workload.sources[0]
    assert (s.radius, s.time, s.flux, s.psi) == (10.0, [0.0], [10.0], [0.1])
    assert s.num_cells > 0


BASE = """%%ListType
Domain:
Lat = 0
Lon = 0
nXtiles = 8
nYtiles = 8
nXpertile = 16
nYpertile = 16
Xtilesize = 16
Boundary Conditions = halt
Cap:
  capX = 0
  capY = 0
  capRadius = 10
  capVolume = 100
{params}
Solver:
T end = 1
Topog:
Type = Function
Topog function = flat
"""


def test_defaults_of_an_almost_empty_deck():
    """No Parameters block: Chezy drag, Mixed erosion with morphodynamics on,
    tanh switch and damping, reference defaults (src/Parameters.f90:41-77,
    486-600); capVolume with capRadius gives the paraboloid height 2V/(pi R^2)
    (src/InitConds.f90:362-368)."""
    run = deck.build(deck.parse(BASE.format(params="")))
    c = run.workload.cfg
    assert c.drag == C.DRAGS["chezy"] and c.chezy_co == 0.01
    assert c.erosion == C.EROSIONS["mixed"] and c.morphodynamics == 1
    assert c.fswitch == C.SWITCHES["tanh"] and c.morpho_damping == C.DAMPINGS["tanh"]
    assert (c.cfl, c.height_threshold, c.tile_buffer) == (0.25, 1e-6, 1)
    assert run.workload.caps[0].height == 2.0 * 100 / math.pi / 10 / 10


def test_damping_resets_without_switch_function():
    """The reference resets MorphoDamp to its default when 'Switch function'
    is absent (src/Parameters.f90:486-490), so a damping given alone has no
    effect; given with a switch function it holds."""
    alone = deck.build(deck.parse(BASE.format(params="Parameters:\nMorphodynamic damping = rat3")))
    assert alone.workload.cfg.morpho_damping == C.DAMPINGS["tanh"]
    both = deck.build(deck.parse(BASE.format(
        params="Parameters:\nSwitch function = step\nMorphodynamic damping = rat3")))
    assert both.workload.cfg.morpho_damping == C.DAMPINGS["rat3"]
    assert both.workload.cfg.fswitch == C.SWITCHES["step"]


def test_erosion_off_and_labels_case_insensitive():
    run = deck.build(deck.parse(BASE.format(
        params="Parameters:\nEROSION = off\ndrag = Pouliquen\nEddy Viscosity = 1e-2")))
    c = run.workload.cfg
    assert c.morphodynamics == 0 and c.erosion == C.EROSIONS["off"]
    assert c.drag == C.DRAGS["pouliquen"] and c.eddy_viscosity == 1e-2
    assert c.diffusive_time_scale == 1.0 / 1e-2


def test_cube_and_periodic_deck_values():
    """Cube_Set (src/InitConds.f90:523-660) and the periodic boundaries of
    the reference's 2D lake-at-rest deck: a level cube over the whole domain
    paints w = 1.5 everywhere (the level surface), and the topography's
    last vertex row and column repeat the first (EqualiseTopographicBoundaryData
    across the periodic edges)."""
    run = deck.load(os.path.join(DECKS, "Input_lake_at_rest_hydro_2d.txt"))
    c = run.workload.cfg
    assert c.bc == C.BC_PERIODIC and (c.nx_tiles, c.ny_tiles) == (1, 1)
    assert np.all(run.workload.init.w == 1.5)
    assert np.array_equal(run.workload.b0v[:, -1], run.workload.b0v[:, 0])
    assert np.array_equal(run.workload.b0v[-1, :], run.workload.b0v[0, :])
    assert np.all(run.workload.init.tile_state == 1)


def test_objects_outside_the_domain_are_refused():
    """The domain checks of Cap_Set, Source_Set and Cube_Set
    (src/InitConds.f90:192-195, 410-413, 643-646); touching the edge is allowed."""
    base = open(os.path.join(DECKS, "Input_lake_at_rest_hydro_2d.txt")).read()
    with pytest.raises(deck.DeckError, match="cube 1 outside"):
        deck.build(deck.parse(base.replace("cubeX = 0.0", "cubeX = 0.5")))
    cap = base.replace("Cube:", "Cap:").replace("cubeX = 0.0", "capX = 10").replace(
        "cubeY = 0.0", "capY = 0").replace("cubeHeight", "capHeight").replace(
        "cubeLength = 40", "capRadius = 10").replace("cubeWidth = 40\n", "").replace(
        "cubeShape = level", "capShape = flat").replace("cubeU", "capU").replace(
        "cubeV", "capV").replace("cubeConc", "capConc")
    deck.build(deck.parse(cap))
    with pytest.raises(deck.DeckError, match="cap 1 outside"):
        deck.build(deck.parse(cap.replace("capX = 10", "capX = 10.5")))


def test_malformed_decks_are_refused():
    with pytest.raises(deck.DeckError):
        deck.build(deck.parse(BASE.format(params="Parameters:\nDrag = Bogus")))
    with pytest.raises(deck.DeckError):
        deck.build(deck.parse(BASE.replace("T end = 1\n", "")))
    with pytest.raises(deck.DeckUnsupported):
        deck.build(deck.parse(BASE.replace("Type = Function", "Type = Raster")))
