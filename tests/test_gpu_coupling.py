"""Coupling truncation on the device (fh_system_create_coupled,
fh_partial_inductance_masked) against fasterhenry/tests/coupling.rs: distant
islands survive truncation within 0.1 %, near ones warn and matter, a
declared pair and the default reproduce the full assembly bit for bit,
truncated entries are exactly 0 and never evaluated, kept ones are the full
fill's bits."""
import math

import numpy as np
import pytest

from coupling_cases import EXTENT, FREQUENCY, SIDE, discretization, two_islands
from fasterhenry_amd import (Coupling, MeshSystem, UnknownGroup, api, model, partial_inductance_matrix,
                             partial_inductance_matrix_masked)
from fasterhenry_amd.errors import SolveError

pytestmark = pytest.mark.gpu


def _impedance(g, ports, coupling, precision="fp64"):
    system = MeshSystem.assemble_with_coupling(g, ports, discretization(), coupling, precision=precision)
    z = system.impedance(FREQUENCY)
    warnings = [str(w) for w in system.truncation_warnings]
    system.close()
    return z, warnings


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(a)


def test_distant_islands_survive_truncation_within_one_part_in_a_thousand():
    g, ports, groups = two_islands(11.0 * EXTENT - SIDE)
    full, _ = _impedance(g, ports, Coupling.all_pairs())
    trunc, warnings = _impedance(g, ports, Coupling.truncated(groups))
    assert warnings == []
    assert _rel(full, trunc) < 1e-3
    for i in range(2):
        assert abs(full[i, i] - trunc[i, i]) / abs(full[i, i]) < 1e-3
    assert full[0, 1] != 0 and trunc[0, 1] == 0 and trunc[1, 0] == 0
    assert abs(full[0, 1]) / abs(full[0, 0]) < 1e-3


def test_near_islands_warn_and_the_dropped_term_is_significant():
    g, ports, groups = two_islands(0.5 * EXTENT)
    full, _ = _impedance(g, ports, Coupling.all_pairs())
    trunc, warnings = _impedance(g, ports, Coupling.truncated(groups))
    assert len(warnings) == 1
    for expected in ("'a'", "'b'", "truncat"):
        assert expected in warnings[0]
    assert abs(full[0, 1]) / abs(full[0, 0]) > 1e-3
    assert _rel(full, trunc) > 1e-3


@pytest.mark.parametrize("precision", ["fp64", "fp32_refined"])
def test_declared_coupling_and_the_default_reproduce_the_full_assembly(precision):
    g, ports, groups = two_islands(0.5 * EXTENT)
    system = MeshSystem.assemble(g, ports, discretization(), precision=precision)
    reference = system.impedance(FREQUENCY)
    system.close()
    default, w = _impedance(g, ports, Coupling.all_pairs(), precision)
    assert np.array_equal(default, reference) and w == []
    declared, w = _impedance(g, ports, Coupling.truncated(groups).coupled("a", "b"), precision)
    assert np.array_equal(declared, reference) and w == []


def test_untagged_segments_share_the_default_group():
    g, ports, groups = two_islands(11.0 * EXTENT - SIDE)
    groups[:4] = [""] * 4
    full, _ = _impedance(g, ports, Coupling.all_pairs())
    trunc, _ = _impedance(g, ports, Coupling.truncated(groups))
    assert trunc[0, 1] == 0 and _rel(full, trunc) < 1e-3


def test_coupling_an_undeclared_group_is_an_error():
    g, ports, groups = two_islands(EXTENT)
    with pytest.raises(UnknownGroup) as e:
        MeshSystem.assemble_with_coupling(g, ports, discretization(),
                                          Coupling.truncated(groups).coupled("a", "typo"))
    assert "typo" in str(e.value)


def test_truncated_pairs_are_never_evaluated_and_kept_ones_are_exact():
    g, ports, groups = two_islands(EXTENT)
    packed = model.pack(g, ports, discretization())
    f = api.discretize_packed(packed)
    assert len(f) == 16
    per = [2] * 8
    full = partial_inductance_matrix(f)
    mask = Coupling.truncated(groups).pair_mask(per)
    assert mask.kept_pairs() == 16 * 17 // 2 - 64
    trunc, methods = partial_inductance_matrix_masked(f, mask, with_methods=True)
    dropped_nonzero = 0
    for i in range(16):
        for j in range(16):
            if mask.keeps(i, j):
                assert trunc[i, j] == full[i, j] and methods[i, j] >= 0
            else:
                assert trunc[i, j] == 0.0 and methods[i, j] == -1  # never evaluated
                dropped_nonzero += full[i, j] != 0.0
    assert dropped_nonzero > 0
    # kernel evaluations = kept pairs: every kept upper-triangle entry has a method
    iu = np.triu_indices(16)
    assert int(np.sum(methods[iu] >= 0)) == mask.kept_pairs()
    declared = partial_inductance_matrix_masked(f, Coupling.truncated(groups).coupled("a", "b").pair_mask(per))
    assert np.array_equal(declared, full)


def test_each_island_is_a_physical_loop():
    g, ports, groups = two_islands(11.0 * EXTENT - SIDE)
    trunc, _ = _impedance(g, ports, Coupling.truncated(groups))
    for i in range(2):
        inductance = trunc[i, i].imag / (2 * math.pi * FREQUENCY)
        assert 10e-9 <= inductance <= 80e-9 and trunc[i, i].real > 0


def test_a_truncation_stays_dense_and_refuses_the_iterative_solver():
    g, ports, groups = two_islands(EXTENT)
    with pytest.raises(SolveError):
        MeshSystem.assemble_with_coupling(g, ports, discretization(), Coupling.truncated(groups),
                                          solver=model.SolverChoice.ITERATIVE)
    system = MeshSystem.assemble_with_coupling(g, ports, discretization(), Coupling.truncated(groups),
                                               solver=model.SolverChoice.AUTO)
    assert system.impedance(1e6).shape == (2, 2)
    system.close()
