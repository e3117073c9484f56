"""The batched sweep (fh_solve_batch): many uneven geometries scheduled
together on the device give, problem by problem, exactly the result of
their own solve; the first failing problem in batch order is reported."""
import numpy as np
import pytest

import oracle
from fasterhenry_amd import MeshError, MeshSystem, fixtures, model, solve_batch
from fasterhenry_amd.model import Port

pytestmark = pytest.mark.gpu

FREQS = [0.0, 1e5, 1e7, 1e9, 3e9]


def _problems():
    b = fixtures.batch_problems()
    return [fixtures.busbar_pair(), fixtures.spiral_system(), fixtures.coupled_structure(),
            b[32], b[9], fixtures.meander_problem(15), b[0]]


@pytest.mark.parametrize("precision", ["fp64", "fp32_refined"])
def test_batch_equals_one_solve_per_problem(precision):
    probs = _problems()
    print(f"batch of {len(probs)} problems, {precision}", flush=True)
    zs, infos = solve_batch(probs, FREQS, precision=precision, concurrency=4)
    for (g, ports, d), z, info in zip(probs, zs, infos):
        packed = model.pack(g, ports, d)
        system = MeshSystem.assemble_packed(packed, ports, precision)
        z1, _ = system.sweep_raw(FREQS)
        system.close()
        assert z.shape == z1.shape
        assert np.array_equal(z.view(np.float64), z1.view(np.float64))
        assert info["filaments"] == packed.filament_count
    assert np.all(zs[0][0].imag == 0.0)


def test_batch_matches_the_oracle():
    probs = _problems()[:3]
    oracle.set_portable_math(True)
    zs, _ = solve_batch(probs, FREQS, precision="fp64", concurrency=3)
    for (g, ports, d), z in zip(probs, zs):
        zr = oracle.System(model.pack(g, ports, d)).sweep(FREQS)
        for k in range(len(FREQS)):
            assert np.linalg.norm(z[k] - zr[k]) / np.linalg.norm(zr[k]) < 1e-12


def test_concurrency_does_not_change_results():
    probs = fixtures.batch_problems(6, seed=7)
    z1, _ = solve_batch(probs, FREQS[1:], concurrency=1)
    z4, _ = solve_batch(probs, FREQS[1:], concurrency=4)
    for a, b in zip(z1, z4):
        assert np.array_equal(a.view(np.float64), b.view(np.float64))


def test_first_failing_problem_is_reported():
    probs = _problems()[:4]
    g, ports, d = fixtures.spiral_system()
    probs[2] = (g, [Port(0, 10_000)], d)  # a port on a node that does not exist
    with pytest.raises(Exception) as e:
        solve_batch(probs, FREQS)
    assert "problem 2" in str(e.value)
