"""Coupling truncation on the host (fasterhenry_amd.coupling) against the
unit tests of fasterhenry/src/coupling.rs:480-677: group resolution, the
pair mask, validation errors and truncation warnings."""
import pytest

from coupling_cases import two_traces
from fasterhenry_amd import (Coupling, NonTransitiveCoupling, SegmentCountMismatch, UnknownGroup)


def test_the_default_couples_everything():
    c = Coupling.all_pairs()
    assert c.is_all_pairs() and c.is_trivial() and c.couples("a", "b")
    assert c.group_of(7) == ""
    mask = c.pair_mask([2, 3])
    assert mask.kept_pairs() == mask.total_pairs()


def test_truncation_drops_exactly_the_cross_group_pairs():
    c = Coupling.truncated(["a", "b"])
    assert not c.is_all_pairs() and c.couples("a", "a") and not c.couples("a", "b")
    mask = c.pair_mask([2, 2])
    assert mask.total_pairs() == 10 and mask.kept_pairs() == 6
    assert mask.keeps(0, 1) and mask.keeps(2, 2)
    assert not mask.keeps(0, 2) and not mask.keeps(3, 1)


def test_declaring_a_pair_keeps_it():
    c = Coupling.truncated(["a", "b", "c"]).coupled("a", "b")
    assert c.couples("a", "b") and c.couples("b", "a") and not c.couples("a", "c")
    mask = c.pair_mask([1, 1, 1])
    assert mask.keeps(0, 1) and not mask.keeps(0, 2) and not mask.keeps(1, 2)
    assert mask.kept_pairs() == 4


def test_untagged_segments_share_the_default_group():
    mask = Coupling.truncated(["", "", "b"]).pair_mask([1, 1, 1])
    assert mask.keeps(0, 1) and not mask.keeps(1, 2)
    short = Coupling.truncated(["b"]).pair_mask([1, 1, 1])
    assert not short.keeps(0, 1) and short.keeps(1, 2)


def test_unknown_and_oversized_declarations_are_errors():
    c = Coupling.truncated(["a", "b"]).coupled("a", "typo")
    with pytest.raises(UnknownGroup) as e:
        c.validate(2)
    assert e.value.name == "typo"
    with pytest.raises(UnknownGroup):
        c.pair_mask([1, 1])
    with pytest.raises(SegmentCountMismatch) as e:
        Coupling.truncated(["a", "b", "c"]).validate(2)
    assert (e.value.expected, e.value.got) == (2, 3)


def test_non_transitive_declarations_are_rejected():
    c = Coupling.truncated(["a", "b", "c"]).coupled("a", "b").coupled("b", "c")
    with pytest.raises(NonTransitiveCoupling) as e:
        c.validate(3)
    assert e.value.groups == ("a", "c")
    Coupling.truncated(["a", "b", "c"]).coupled("a", "b").coupled("b", "c").coupled("a", "c").validate(3)
    Coupling.truncated(["a", "b", "c"]).coupled("a", "b").validate(3)


def test_near_truncated_groups_warn_and_distant_ones_do_not():
    c = Coupling.truncated(["a", "b"])
    near = c.truncation_warnings(two_traces(0.25))
    assert len(near) == 1 and near[0].groups == ("a", "b")
    assert abs(near[0].separation - 0.25) < 1e-12 and abs(near[0].extent - 1.0) < 1e-12
    assert "'a'" in str(near[0]) and "truncated" in str(near[0])
    assert str(near[0]) == ("mutual inductance between groups 'a' and 'b' is truncated, but they are only "
                            "2.500e-1 m apart — 0.25x their 1.000e0 m extent; truncation is justified "
                            "from about 10x")
    assert c.truncation_warnings(two_traces(4.0)) == []
    assert Coupling.all_pairs().truncation_warnings(two_traces(0.01)) == []
    assert c.coupled("a", "b").truncation_warnings(two_traces(0.01)) == []


def test_the_default_group_is_named_in_warnings():
    w = Coupling.truncated(["", "b"]).truncation_warnings(two_traces(0.1))
    assert len(w) == 1 and "(default)" in str(w[0])
