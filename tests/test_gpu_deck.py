"""Decks end to end on the device: the reference's fixture decks read by
fh_deck_parse and solved as the CLI solves them (deck.run), against the
oracle's solve of the same problem."""
import os

import numpy as np
import pytest

import oracle
from fasterhenry_amd import deck, model

pytestmark = pytest.mark.gpu

DECKS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "decks")


@pytest.mark.parametrize("name", ["spiral.inp", "twobar.inp"])
@pytest.mark.parametrize("precision", ["fp64", "fp32_refined"])
def test_decks_solve_as_the_oracle(name, precision):
    oracle.set_portable_math(True)
    d = deck.read(os.path.join(DECKS, name))
    result = deck.run(d, precision=precision)
    z_ref = oracle.solve(model.pack(d.geometry, d.ports, d.discretization), d.frequencies)
    for k in range(len(d.frequencies)):
        assert np.linalg.norm(result.impedance_ohm[k] - z_ref[k]) / np.linalg.norm(z_ref[k]) < 1e-10
    assert [p.name for p in result.ports] == [p.name for p in d.ports]


def test_a_couples_deck_truncates_the_mutual():
    with open(os.path.join(DECKS, "twobar.inp")) as fh:
        text = fh.read()
    text = text.replace("e1 na nb", "e1 na nb group=A").replace("e2 nc nd", "e2 nc nd group=B")
    text = text.replace(".end", ".couples A\n.end")
    d = deck.parse(text)
    assert not d.coupling.is_all_pairs()
    z = deck.run(d, [1e6], precision="fp64").impedance_ohm[0]
    assert z[0, 1] == 0 and z[1, 0] == 0
    full = deck.run(deck.parse(text.replace(".couples A\n", "")), [1e6], precision="fp64").impedance_ohm[0]
    assert full[0, 1] != 0 and abs(z[0, 0] - full[0, 0]) / abs(full[0, 0]) < 1e-2
