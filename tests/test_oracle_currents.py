"""The oracle's filament currents (oracle.currents) against the known
answers of fasterhenry/tests/currents.rs: uniform |J| at DC, the skin-effect
profile of a wide slab, power balance, Kirchhoff's current law on a ring,
ports without internal loops."""
import numpy as np

import oracle
from currents_cases import COPPER, bar, ring, slab_frequency, slab_profile, two_bars
from fasterhenry_amd import model


def test_bar_at_dc_has_uniform_density():
    g, ports = bar(10e-3, 1e-3, 35e-6)
    packed = model.pack(g, ports, model.Discretization.uniform(5, 3))
    area = 1e-3 * 35e-6
    sol = oracle.currents(packed, 0.0, [2.0])
    assert len(sol["current"]) == 15
    assert np.all(np.abs(sol["density"] - 2.0 / area) < 1e-6 * 2.0 / area)
    assert np.all(np.abs(sol["current"].imag) < 1e-12)
    assert abs(sol["segment_current"][0] - 2.0) < 1e-9
    assert len(sol["peaks"]) == 1
    r = 10e-3 / (COPPER * area)
    assert abs(sol["dissipated_w"] - 0.5 * 4.0 * r) < 1e-9 * r
    assert sol["relative_residual"] < 1e-9


def test_wide_slab_follows_the_skin_effect_profile():
    length, width, t = 100e-3, 40e-3, 200e-6
    delta = t / 4.0
    f = slab_frequency(delta)
    assert abs(model.skin_depth(f, COPPER) - delta) < 1e-12
    g, ports = bar(length, width, t)
    packed = model.pack(g, ports, model.Discretization.uniform(1, 40))
    sol = oracle.currents(packed, f, [1.0])
    profile, surface = slab_profile(width, t, delta)
    worst = max(abs(d - profile(loc[2])) for d, loc in zip(sol["density"], sol["location"]))
    assert worst / surface < 0.05
    k = sol["peaks"][0]
    assert abs(sol["location"][k][2]) > 0.45 * t
    assert abs(sol["density"][k] - profile(sol["location"][k][2])) < 0.05 * surface
    assert sol["density"][k] > 2.0 / (width * t)
    assert sol["relative_residual"] < 1e-8


def test_power_balance_and_kcl_on_a_ring():
    g, ports = ring()
    packed = model.pack(g, ports, model.Discretization.uniform(3, 2))
    system = oracle.System(packed)
    for f in [0.0, 1e5, 1e8]:
        sol = oracle.currents(packed, f, [1.0], system)
        assert sol["relative_residual"] < 1e-9
        # the direct side plus the way round (side 1 traversed backwards)
        # carry the port current together
        assert abs(sol["segment_current"][0] - sol["segment_current"][1] - 1.0) < 1e-9
        assert len(sol["segment_current"]) == 4 and len(sol["peaks"]) == 1
    z = system.sweep([1e8])[0][0, 0]
    sol = oracle.currents(packed, 1e8, [1.0], system)
    assert abs(sol["input_w"] - 0.5 * z.real) < 1e-9 * z.real


def test_ports_without_internal_loops():
    g, ports = two_bars()
    packed = model.pack(g, ports, model.Discretization.uniform(1, 1))
    ex = [1.0, 0.5j]
    sol = oracle.currents(packed, 1e6, ex)
    assert abs(sol["current"][0] - ex[0]) < 1e-12 and abs(sol["current"][1] - ex[1]) < 1e-12
    assert len(sol["peaks"]) == 2
    assert sol["relative_residual"] < 1e-9
