"""The refined form's escalation (fh_system_set_refinement): split-bf16
factor with classical refinement, GMRES-IR on the same factor, the FP32
factor (classical, then GMRES-IR), and a plain FP64 solve of whatever is
left, so that a frequency the reference's LU solves (it fails only on an
exactly zero pivot, dense.rs:43-84) is not refused.

The ill-conditioned cases are graded busbar pairs whose surface cells are
thousands of times thinner than their middle ones (12 x 6 cells at ratio 3,
20 x 6 at ratio 5): the loop resistances span that range, and numpy measures
condition numbers of Zee from 4e5 to 7e9.  Each stage is also run alone on a
well-conditioned problem and checked against the plain FP64 form.
"""
import numpy as np
import pytest

import oracle
from fasterhenry_amd import (MeshSystem, NotConverged, Singular, fixtures, model, native)

FREQS = [0.0, 1e3, 1e6, 1e8, 1e9, 1e10]


def graded_pair(nw, nh, ratio, length=1e-3, width=100e-6, thickness=35e-6, gap=1e-6, segs=4):
    """Two bars `gap` apart edge to edge, shorted at the far end, one port
    across the near ends; every segment nw x nh graded by `ratio`."""
    g = model.Geometry()
    a = [g.add_node(model.Node(length * k / segs, 0.0, 0.0)) for k in range(segs + 1)]
    b = [g.add_node(model.Node(length * k / segs, width + gap, 0.0)) for k in range(segs + 1)]
    for k in range(segs):
        g.add_segment(model.SegmentDef(a[k], a[k + 1], width, thickness, fixtures.COPPER))
    for k in range(segs):
        g.add_segment(model.SegmentDef(b[k + 1], b[k], width, thickness, fixtures.COPPER))
    g.add_segment(model.SegmentDef(a[-1], b[-1], width, thickness, fixtures.COPPER))
    return g, [model.Port(a[0], b[0])], model.Discretization.graded(nw, nh, ratio)


@pytest.fixture(autouse=True)
def _portable():
    oracle.set_portable_math(True)
    yield


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_graded_pair_is_ill_conditioned():
    """numpy's condition number of Zee = R + jwL from the oracle's blocks."""
    s = oracle.System(model.pack(*graded_pair(12, 6, 3.0)))
    b = s.blocks()
    kappa = np.linalg.cond(b["r_ee"])  # DC
    assert kappa > 1e5


@pytest.mark.gpu
def test_ill_conditioned_sweep_escalates_and_matches_the_oracle():
    prob = graded_pair(12, 6, 3.0)
    packed = model.pack(*prob)
    zr = oracle.System(packed).sweep(FREQS)
    system = MeshSystem.assemble_packed(packed, prob[1], "fp32_refined")
    res = system.sweep(FREQS)
    err = max(rel(res.impedance_ohm[k], zr[k]) for k in range(len(FREQS)))
    print("stage", res.extra["refine_stage"], "arith", res.extra["factor_arith"], "err", err, flush=True)
    # split-bf16 classical refinement alone does not converge here
    assert res.extra["refine_stage"] >= 1
    assert err <= 1e-10
    # and with that stage alone the sweep is refused, by frequency
    system.set_refinement(native.FH_REFINE_SPLIT_CLASSICAL)
    with pytest.raises(NotConverged) as e:
        system.sweep(FREQS)
    assert 0 <= e.value.index0 < len(FREQS)
    system.close()


@pytest.mark.gpu
def test_very_ill_conditioned_sweep_is_not_refused():
    """kappa(Zee) ~ 7e9: every FP32 stage may fail; the FP64 stage solves it,
    to the conditioning's own accuracy against the oracle's pivoted LU."""
    prob = graded_pair(20, 6, 5.0, width=200e-6)
    packed = model.pack(*prob)
    zr = oracle.System(packed).sweep(FREQS)
    system = MeshSystem.assemble_packed(packed, prob[1], "fp32_refined")
    res = system.sweep(FREQS)
    z64 = MeshSystem.assemble_packed(packed, prob[1], "fp64").sweep(FREQS).impedance_ohm
    system.close()
    print("stage", res.extra["refine_stage"], flush=True)
    for k in range(len(FREQS)):
        assert np.all(np.isfinite(res.impedance_ohm[k]))
        # both FP64 methods agree to kappa * eps; the refined Z to the FP64 GPU Z
        assert rel(z64[k], zr[k]) <= 1e-5
        assert rel(res.impedance_ohm[k], z64[k]) <= 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("stages,stage,arith", [
    (native.FH_REFINE_SPLIT_CLASSICAL, 0, 2),
    (native.FH_REFINE_SPLIT_GMRES, 1, 2),
    (native.FH_REFINE_FP32_CLASSICAL, 2, 1),
    (native.FH_REFINE_FP32_GMRES, 3, 1),
    (native.FH_REFINE_FP64, 4, 0),
])
def test_each_stage_alone_matches_the_fp64_form(stages, stage, arith):
    prob = fixtures.meander_problem(15)
    packed = model.pack(*prob)
    freqs = [0.0] + list(fixtures.log_frequencies(1e4, 1e9, 9))
    z64 = MeshSystem.assemble_packed(packed, prob[1], "fp64").sweep(freqs).impedance_ohm
    system = MeshSystem.assemble_packed(packed, prob[1], "fp32_refined")
    system.set_refinement(stages)
    res = system.sweep(freqs)
    system.close()
    assert res.extra["refine_stage"] == stage
    if stage < 4:
        assert res.extra["factor_arith"] == arith
    err = max(rel(res.impedance_ohm[k], z64[k]) for k in range(len(freqs)))
    assert err <= (1e-13 if stage == 4 else 1e-10), err
    assert np.all(res.impedance_ohm[0].imag == 0.0)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32_refined"])
def test_non_finite_impedance_is_singular_at_its_frequency(precision):
    """A non-finite Z is SolveError::Singular at the first such frequency
    (solve.rs:806-830, 1055-1060), as the oracle reports it."""
    prob = fixtures.spiral_system()
    packed = model.pack(*prob)
    freqs = [1e3, 1e6, 1e308]
    with pytest.raises(oracle.OracleError) as eo:
        oracle.System(packed).sweep(freqs)
    assert eo.value.status == oracle.FHO_E_SINGULAR
    system = MeshSystem.assemble_packed(packed, prob[1], precision)
    with pytest.raises(Singular) as e:
        system.sweep(freqs)
    assert e.value.index0 == 2 and e.value.value == 1e308
    # the system is still usable afterwards
    z = system.sweep(freqs[:2]).impedance_ohm
    assert np.all(np.isfinite(z))
    system.close()


def test_refinement_mask_is_validated():
    """Host-side argument check (no system needed for the null case)."""
    L = native.lib()
    assert L.fh_system_set_refinement(None, native.FH_REFINE_ALL) != native.FH_OK
