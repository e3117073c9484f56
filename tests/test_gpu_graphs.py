"""Graph replay of repeated refined sweeps (fh_system_set_graphs): from the
second sweep of a chunk with the same frequency count, each sweep part's
split factorization and first two classical refinement steps, with their
acceptance test on the device, run as one captured HIP graph.  Replayed
sweeps must give the same bits as eager ones, on every schedule, when the
frequency count changes (a new capture) and changes back, and after the
frequencies themselves change (the same graph, new inputs).
"""
import numpy as np
import pytest

from fasterhenry_amd import MeshSystem, fixtures, model

pytestmark = pytest.mark.gpu

FREQS = [0.0] + list(fixtures.log_frequencies(1e4, 1e9, 11))
OTHER = list(fixtures.log_frequencies(2e4, 3e9, 12))


@pytest.fixture(scope="module")
def problem():
    prob = fixtures.meander_problem(15)
    return model.pack(*prob), prob[1]


def bits(z):
    return np.ascontiguousarray(z).view(np.uint64)


def eager(packed, ports, freqs, schedule):
    s = MeshSystem.assemble_packed(packed, ports, "fp32_refined")
    s.set_graphs(False)
    s.set_schedule(*schedule)
    out = [s.sweep_raw(f) for f in freqs]
    s.close()
    return out


@pytest.mark.parametrize("schedule", [(256, 1), (256, 2), (128, 3)])
def test_replayed_sweeps_match_eager_ones_bit_for_bit(problem, schedule):
    packed, ports = problem
    order = [FREQS, FREQS, FREQS, OTHER, FREQS[:6], FREQS[:6], FREQS[:6], FREQS]
    ref = eager(packed, ports, order, schedule)
    s = MeshSystem.assemble_packed(packed, ports, "fp32_refined")
    s.set_schedule(*schedule)
    for k, f in enumerate(order):
        z, info = s.sweep_raw(f)
        zr, ir = ref[k]
        print("schedule", schedule, "sweep", k, "steps", info.refine_steps_max, flush=True)
        assert np.array_equal(bits(z), bits(zr)), (schedule, k)
        assert info.refine_steps_max == ir.refine_steps_max
        assert info.refine_stage == ir.refine_stage
        assert info.refine_change_max == ir.refine_change_max
    s.close()


def test_profiled_replays_count_the_captured_updates(problem):
    """With profiling on, a replayed sweep reports the same number of timed
    update launches and the same flops as an eager one, and a positive time."""
    packed, ports = problem
    s = MeshSystem.assemble_packed(packed, ports, "fp32_refined")
    s.set_schedule(256, 1)
    s.set_profiling(True)
    _, first = s.sweep_raw(FREQS)
    for _ in range(3):
        z, info = s.sweep_raw(FREQS)
        assert info.update_kernel_launches == first.update_kernel_launches
        assert info.update_kernel_flops == first.update_kernel_flops
        assert info.update_kernel_ms > 0.0
    s.set_graphs(False)
    z2, _ = s.sweep_raw(FREQS)
    assert np.array_equal(bits(z), bits(z2))
    s.close()
