"""The `.inp` deck reader (fh_deck_parse; host code, no GPU) against the
reference's fixture decks (fasterhenry-cli/tests/data, copied under
tests/golden/decks) and the rules of fasterhenry-cli/src/inp.rs."""
import os

import numpy as np
import pytest

from fasterhenry_amd import DeckError, deck, model

HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = os.path.join(HERE, "golden", "decks")


def _read(name):
    with open(os.path.join(DECKS, name)) as fh:
        return fh.read()


def _packed(d):
    return model.pack(d.geometry, d.ports, d.discretization)


def test_the_spiral_deck_is_the_spiral_json_problem():
    d = deck.parse(_read("spiral.inp"))
    j = deck.problem_from_json(_read("spiral.json"))
    assert d.title == "2-turn square spiral"
    a, b = _packed(d), _packed(j)
    assert np.array_equal(a.nodes, b.nodes)
    assert a.segments.tobytes() == b.segments.tobytes()
    assert a.grids.tobytes() == b.grids.tobytes()
    assert a.ports.tobytes() == b.ports.tobytes()
    assert [p.name for p in d.ports] == ["spiral"] == [p.name for p in j.ports]
    assert d.frequencies == j.frequencies == [1e6, 1e7, 1e8, 1e9]
    assert d.discretization.kind == "uniform" and d.coupling.is_all_pairs()
    # and the JSON form round-trips
    again = deck.problem_from_json(deck.problem_to_json(j))
    assert _packed(again).segments.tobytes() == b.segments.tobytes()


def test_the_two_bar_deck():
    d = deck.parse(_read("twobar.inp"))
    assert len(d.geometry.nodes) == 4 and len(d.geometry.segments) == 2
    assert [(p.positive, p.negative, p.name) for p in d.ports] == [(0, 1, "barA"), (2, 3, "barB")]
    assert d.geometry.nodes[2] == model.Node(0.0, 5e-4, 0.0)


BASE = """.units mm
.default z=0 w=0.2 h=0.035 sigma=5.8e4
N1 x=0 y=0
N2 x=1 y=0
E1 N1 N2
.external N1 N2
.freq fmin=1e3 fmax=1e7 ndec=0.5
.end
"""


def test_units_scale_lengths_and_conductivity():
    d = deck.parse(BASE)
    s = d.geometry.segments[0]
    assert s.width == 0.2 * 1e-3 and s.height == 0.035 * 1e-3
    assert s.sigma == 5.8e4 / 1e-3
    assert d.geometry.nodes[1].x == 1e-3
    rho = deck.parse(BASE.replace("sigma=5.8e4", "rho=1.7241e-5")).geometry.segments[0].sigma
    assert rho == 1.0 / 1.7241e-5 / 1e-3
    assert d.frequencies == [1e3, 1e5, 1e7]  # ndec need not be an integer
    assert d.ports[0].name == "n1/n2"  # default label, node names lowercased


def test_decade_sweep_rule():
    d = deck.parse(BASE.replace("fmin=1e3 fmax=1e7 ndec=0.5", "fmin=1 fmax=1000 ndec=3"))
    assert d.frequencies[0] == 1.0 and d.frequencies[3] == 10.0 and d.frequencies[-1] == 1000.0
    assert len(d.frequencies) == 10
    assert abs(d.frequencies[1] - 10 ** (1 / 3)) < 1e-15
    assert deck.parse(BASE.replace("fmin=1e3 fmax=1e7", "fmin=0 fmax=0")).frequencies == [0.0]


def test_continuations_spacing_and_case():
    text = BASE.replace("E1 N1 N2", "e1 n1\n+ N2 nwinc = 3\n+ nhinc= 2 rw =2")
    d = deck.parse(text)
    assert d.discretization.kind == "per_segment_graded"
    g = d.discretization.value[0]
    assert (g.nw, g.nh, g.width_ratio, g.height_ratio) == (3, 2, 2.0, 1.0)
    # without a ratio the grid is uniform
    d = deck.parse(BASE.replace("E1 N1 N2", "E1 N1 N2 nwinc=3 nhinc=2"))
    assert d.discretization.kind == "uniform" and d.discretization.value == model.Subdivision(3, 2)


def test_equiv_joins_nodes_and_unreferenced_nodes_are_dropped():
    text = BASE.replace("N2 x=1 y=0\n", "N2 x=1 y=0\nN3 x=1 y=0\nN9 x=5 y=5\n.equiv N2 N3 Nalias\n") \
               .replace("E1 N1 N2", "E1 N1 N3\nE2 N3 Nalias w=0.1")
    with pytest.raises(DeckError):
        deck.parse(text)  # E2 from the node to itself: zero length
    text = BASE.replace("N2 x=1 y=0\n", "N2 x=1 y=0\nN3 x=2 y=0\nN9 x=5 y=5\n.equiv N3 Nalias\n") \
               .replace("E1 N1 N2", "E1 N1 N2\nE2 N2 Nalias")
    d = deck.parse(text)
    assert len(d.geometry.nodes) == 3  # N9 is never referenced
    assert d.geometry.segments[1].b == 2


@pytest.mark.parametrize("change,line", [
    (lambda t: t.replace(".units mm\n", ""), None),
    (lambda t: t.replace(".end\n", ""), None),
    (lambda t: t + "N7 x=0 y=0\n", 9),
    (lambda t: t.replace("E1 N1 N2", "E1 N1 N5"), 5),
    (lambda t: t.replace("N2 x=1", "n1 x=1"), 4),
    (lambda t: t.replace("E1 N1 N2", "E1 N1 N2 sigma=1 rho=1"), 5),
    (lambda t: t.replace("E1 N1 N2", "G1 x1=0"), 5),
    (lambda t: t.replace("E1 N1 N2", "E1 N1 N2 rw=0.5"), 5),
    (lambda t: t.replace(".units mm", ".units furlong"), 1),
    (lambda t: t.replace("ndec=0.5", "ndec=0"), 7),
    (lambda t: "+ continued\n" + t, 1),
    (lambda t: t.replace("E1 N1 N2", "E1 N1 N2 nwinc=4294967297"), 5),
    (lambda t: t.replace("E1 N1 N2", "E1 N1 N2 nhinc=18446744073709551615"), 5),
    (lambda t: t.replace("E1 N1 N2", "E1 N1 N2 nwinc=123456789012345678901"), 5),
    (lambda t: t.replace("E1 N1 N2", "E1 N1 N2 nwinc=0"), 5),
    (lambda t: t.replace("E1 N1 N2", "E1 N1 N2 nwinc=2.5"), 5),
])
def test_rejected_decks_name_their_line(change, line):
    """Counts beyond a segment grid's 32 bits are refused by line instead of
    wrapping (nwinc = 2^32 + 1 must not become 1 filament)."""
    with pytest.raises(DeckError) as e:
        deck.parse(change(BASE))
    if line is not None:
        assert e.value.index0 == line, str(e.value)
        assert str(e.value).startswith(f"line {line}:")


def test_couples_builds_a_truncating_coupling():
    text = BASE.replace("E1 N1 N2", "N3 x=0 y=5\nN4 x=1 y=5\nE1 N1 N2 group=a\nE2 N3 N4 group=b") \
               .replace(".end", ".external N3 N4\n.couples a\n.end")
    d = deck.parse(text)
    assert not d.coupling.is_all_pairs()
    assert d.coupling.segment_groups == ["a", "b"] and not d.coupling.couples("a", "b")
    d = deck.parse(text.replace(".couples a", ".couples a b"))
    assert d.coupling.couples("a", "b")
    with pytest.raises(DeckError):
        deck.parse(text.replace(".couples a", ".couples a typo"))
    with pytest.raises(DeckError):
        deck.parse(text.replace(".couples a", ".couples all\n.couples a"))
