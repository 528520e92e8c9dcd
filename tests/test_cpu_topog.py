"""The TopogFunc family (src/TopogFuncs.f90) of the deck reader against the
reference's own compiled subroutines: the committed vectors
(tests/golden/make_topog_vectors.py) and, where oracle/_ref is built, fresh
evaluations on another grid.  Bit for bit."""
import os

import numpy as np
import pytest

from kestrel_amd import deck
from oracle.oracle import REF_TOPOG, ref_available, ref_topog

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
