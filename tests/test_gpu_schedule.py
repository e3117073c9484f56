"""The sweep schedule and its autotuner (fh_system_set_schedule,
fh_system_autotune, fh_tune_save / fh_tune_load): every candidate schedule
gives the same Z to the refinement's accuracy, the tuner picks one of the
candidates and stores it for the size class, a new system of that class
takes it, and the cache survives a JSON round trip."""
import json
import os

import numpy as np
import pytest

import oracle
from fasterhenry_amd import InvalidArgument, MeshSystem, fixtures, model, tune_load, tune_save

pytestmark = pytest.mark.gpu

FREQS = [0.0] + list(fixtures.log_frequencies(1e4, 1e9, 11))


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def problem():
    prob = fixtures.meander_problem(15)
    return model.pack(*prob), prob[1]


def test_every_schedule_gives_the_same_impedance(problem):
    packed, ports = problem
    z64 = MeshSystem.assemble_packed(packed, ports, "fp64").sweep(FREQS).impedance_ohm
    system = MeshSystem.assemble_packed(packed, ports, "fp32_refined")
    assert system.schedule() == {"outer_block": 256, "sweep_parts": 2}
    for nb in (64, 128, 256, 512):
        for parts in (1, 2, 3):
            system.set_schedule(nb, parts)
            assert system.schedule() == {"outer_block": nb, "sweep_parts": parts}
            z = system.sweep(FREQS).impedance_ohm
            err = max(rel(z[k], z64[k]) for k in range(len(FREQS)))
            assert err <= 1e-10, (nb, parts, err)
    for bad in ((96, 1), (0, 1), (2048, 1), (256, 0), (256, 5)):
        with pytest.raises(InvalidArgument):
            system.set_schedule(*bad)
    system.close()
    fp64 = MeshSystem.assemble_packed(packed, ports, "fp64")
    zo = oracle.System(packed).sweep(FREQS)
    for nb in (64, 256):
        fp64.set_schedule(nb, 1)
        z = fp64.sweep(FREQS).impedance_ohm
        assert max(rel(z[k], zo[k]) for k in range(len(FREQS))) <= 1e-12
    fp64.close()


def test_autotune_picks_a_candidate_and_the_cache_carries_it(problem, tmp_path):
    packed, ports = problem
    system = MeshSystem.assemble_packed(packed, ports, "fp32_refined")
    chosen, ms = system.autotune(FREQS)
    assert chosen["outer_block"] in (128, 256, 512) and chosen["sweep_parts"] in (1, 2, 3)
    assert ms > 0.0 and system.schedule() == chosen
    z = system.sweep(FREQS).impedance_ohm
    system.close()
    # a new system of the same size class takes the tuned schedule at its sweep
    other = MeshSystem.assemble_packed(packed, ports, "fp32_refined")
    z2 = other.sweep(FREQS).impedance_ohm
    assert other.schedule() == chosen
    assert np.array_equal(z.view(np.float64), z2.view(np.float64))
    other.close()
    path = os.path.join(tmp_path, "tune.json")
    tune_save(path)
    doc = json.load(open(path))
    assert doc["version"] == 1
    keys = {e["key"]: e for e in doc["entries"]}
    m = 1 << (MeshSystem.assemble_packed(packed, ports).counts().internal_meshes - 1).bit_length()
    entry = keys[f"dense/refined/m{m}/k16"]
    assert entry["outer_block"] == chosen["outer_block"] and entry["sweep_parts"] == chosen["sweep_parts"]
    tune_load(path)
    with pytest.raises(InvalidArgument):
        tune_load(os.path.join(tmp_path, "missing.json"))
    # an FP64 system tunes its outer block
    fp64 = MeshSystem.assemble_packed(packed, ports, "fp64")
    c64, _ = fp64.autotune(FREQS)
    assert c64["outer_block"] in (64, 128, 256) and c64["sweep_parts"] == 1
    fp64.close()
