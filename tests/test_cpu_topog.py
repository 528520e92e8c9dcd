"""The TopogFunc family (src/TopogFuncs.f90) of the deck reader and the
bicubic resampling of elevation data (src/Interp2d.f90) against the
reference's own compiled routines: the committed vectors
(tests/golden/make_topog_vectors.py) and, where oracle/_ref is built, fresh
evaluations.  Bit for bit; the bicubic also against closed forms."""
import os

import numpy as np
import pytest

from kestrel_amd import deck, dem
from oracle.oracle import REF_TOPOG, ref_available, ref_bicubic, ref_topog

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "topog_funcs.npz"))


@pytest.mark.parametrize("name", REF_TOPOG)
def test_against_golden_vectors(name):
    key = name.replace(" ", "_")
    got = deck.topog_function(name, GOLD[f"{key}__params"], GOLD["xv"], GOLD["yv"])
    assert np.array_equal(got, GOLD[f"{key}__b0"])


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("name", REF_TOPOG)
def test_against_the_reference_on_a_deck_grid(name):
    from kestrel_amd import config as C, scenario as S
    cfg = C.default_config(6, 4, 16, 16, dx=2.5)
    C.derive(cfg)
    xv, yv = S.vertex_coords(cfg)
    rng = np.random.default_rng(17 + REF_TOPOG.index(name))
    key = name.replace(" ", "_")
    p = GOLD[f"{key}__params"] * rng.uniform(0.8, 1.2, len(GOLD[f"{key}__params"]))
    assert np.array_equal(deck.topog_function(name, p, xv, yv), ref_topog(name, p, xv, yv))


def test_deck_parameter_count_is_checked():
    with pytest.raises(deck.DeckError):
        d = deck.Deck.__new__(deck.Deck)
        d.topog = {"type": "Function", "topog function": "xtrislope", "topog params": "(1, 2, 3)"}
        from kestrel_amd import config as C
        c = C.default_config(4, 4, 8, 8)
        deck._topography(d, c)


def test_bicubic_against_golden_vectors():
    got = dem.bicubic(GOLD["bicubic_z"], GOLD["bicubic_x"], GOLD["bicubic_y"])
    assert np.array_equal(got, GOLD["bicubic_out"])


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built")
def test_bicubic_against_the_reference():
    rng = np.random.default_rng(5)
    z = rng.uniform(-50.0, 500.0, (40, 30))
    x = np.concatenate([rng.uniform(2, 39, 2000), np.arange(2, 39).astype(float)])
    y = np.concatenate([rng.uniform(2, 29, 2000), np.full(37, 7.0)])
    assert np.array_equal(dem.bicubic(z, x, y), ref_bicubic(z, x, y))


def test_bicubic_closed_forms():
    """Exact at the data nodes; a bilinear-plus-quadratic surface (which the
    central differences and the cubic Hermite patch reproduce) to rounding;
    the coefficient matrix is integer and inverts the corner map."""
    i = np.arange(1, 31, dtype=float)[:, None]
    j = np.arange(1, 21, dtype=float)[None, :]
    z = 3.0 + 2.0 * i - j + 0.5 * i * j + 0.25 * i * i - 0.125 * j * j
    x = np.array([5.0, 7.25, 12.5, 20.75, 28.0])
    y = np.array([4.0, 9.5, 3.25, 17.0, 12.75])
    exact = 3.0 + 2.0 * x - y + 0.5 * x * y + 0.25 * x * x - 0.125 * y * y
    assert np.allclose(dem.bicubic(z, x, y), exact, rtol=0, atol=1e-12)
    assert np.array_equal(dem.bicubic(z, np.array([10.0]), np.array([6.0])), z[9:10, 5])
    assert np.array_equal(dem.M, np.rint(dem.M)) and abs(np.linalg.det(dem.M)) > 0


def test_bicubic_rejects_points_without_margin():
    z = np.zeros((10, 10))
    with pytest.raises(ValueError):
        dem.bicubic(z, np.array([1.5]), np.array([5.0]))


def test_tile_heights_four_point_mean():
    """TileHeightData's vertex value (src/dem.f90:320-340): the mean of the
    bicubic at the four points vertex +- dx/2, +- dy/2 in data coordinates.
    A surface the bicubic reproduces (bilinear plus quadratic) gives the
    four-point mean of its exact values; on a random raster each of the four
    evaluations is the reference's compiled Bicubic when that is built."""
    i = np.arange(1, 41, dtype=float)[:, None]
    j = np.arange(1, 31, dtype=float)[None, :]
    z = 3.0 + 2.0 * i - j + 0.5 * i * j + 0.25 * i * i - 0.125 * j * j
    xv = np.array([-4.0, -1.5, 0.0, 2.5, 6.0])
    yv = np.array([-3.0, 0.0, 1.25, 4.0])
    centre, origin, pixel, dx, dy = (500.0, 900.0), (480.0, 880.0), (1.0, 1.0), 2.0, 1.0
    got = dem.tile_heights(z, xv, yv, dx, dy, centre, origin, pixel)
    X, Y = np.meshgrid(xv, yv, indexing="ij")

    def f(px, py):
        x = (px - origin[0]) / pixel[0] + 1
        y = (py - origin[1]) / pixel[1] + 1
        return 3.0 + 2.0 * x - y + 0.5 * x * y + 0.25 * x * x - 0.125 * y * y

    want = 0.25 * sum(f(centre[0] + X + sx * dx / 2, centre[1] + Y + sy * dy / 2)
                      for sx in (-1, 1) for sy in (-1, 1))
    assert np.allclose(got, want, rtol=0, atol=1e-11)
    # a planted mistake (the vertex itself, no offsets) differs on a curved surface
    assert not np.allclose(dem.tile_heights(z, xv, yv, 0.0, 0.0, centre, origin, pixel), want, atol=1e-6)
    if ref_available():
        rng = np.random.default_rng(11)
        zr = rng.uniform(-50.0, 500.0, (40, 30))
        got = dem.tile_heights(zr, xv, yv, dx, dy, centre, origin, pixel)
        h = [ref_bicubic(zr, ((centre[0] + X + (a - 0.5) * dx) - origin[0]).ravel() / pixel[0] + 1,
                         ((centre[1] + Y + (b - 0.5) * dy) - origin[1]).ravel() / pixel[1] + 1).reshape(X.shape)
             for a in (0, 1) for b in (0, 1)]
        assert np.array_equal(got, 0.25 * (((h[0] + h[1]) + h[2]) + h[3]))
