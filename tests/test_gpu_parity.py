"""GPU parity against the oracle, through the C ABI.

* device log/atan: bit for bit against the oracle's portable math;
* the fill (partial_inductance_matrix, batch.rs:246-253): every entry and
  every method bit for bit on fixtures that exercise all five methods;
* mesh blocks (congruence, solve.rs:1172-1200): bit for bit;
* Z(w): FP64 plain path within 1e-12 relative (Frobenius per frequency, the
  reference's metric, tests/solve.rs:114-116) of the oracle's LU + Schur;
* DC in real arithmetic, Im Z exactly 0 (solve.rs:809-812);
* error reporting: invalid frequency index, kernel errors.
"""
import numpy as np
import pytest

import oracle
from fasterhenry_amd import (InvalidArgument, InvalidFrequency, MeshSystem, fixtures, model,
                             partial_inductance_matrix, device_math)

pytestmark = pytest.mark.gpu

Z_TOL_FP64 = 1e-12


@pytest.fixture(autouse=True)
def _portable():
    oracle.set_portable_math(True)
    yield


def _fixtures():
    return {
        "busbar": fixtures.busbar_pair(),
        "spiral": fixtures.spiral_system(),
        "coupled": fixtures.coupled_structure(),
        "meander15": fixtures.meander_problem(15),
        "cloud_manhattan": fixtures.random_cloud_problem(200, True, 7),
        "cloud_skew": fixtures.random_cloud_problem(200, False, 7),
    }


def rel_frobenius(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_device_math_matches_portable_bits():
    rng = np.random.default_rng(1)
    x = np.concatenate([
        np.exp(rng.uniform(-740, 709, 20000)),
        rng.uniform(-8, 8, 20000),
        1.0 / rng.uniform(-3, 3, 5000),
        [0.0, -0.0, 1.0, -1.0, 0.4375, 0.6875, 1.1875, 2.4375, 2.0 ** -1074, 1e308,
         np.inf, -np.inf],
    ])
    lg, at = device_math(x)
    want_lg = np.array([oracle.log(v) for v in x])
    want_at = np.array([oracle.atan(v) for v in x])
    same_lg = (lg.view(np.uint64) == want_lg.view(np.uint64)) | (np.isnan(lg) & np.isnan(want_lg))
    same_at = (at.view(np.uint64) == want_at.view(np.uint64)) | (np.isnan(at) & np.isnan(want_at))
    assert same_lg.all(), x[~same_lg][:5]
    assert same_at.all(), x[~same_at][:5]


@pytest.mark.parametrize("name", list(_fixtures().keys()))
def test_fill_is_bit_identical(name):
    g, ports, d = _fixtures()[name]
    packed = model.pack(g, ports, d)
    f = oracle.discretize(packed)
    L, methods = partial_inductance_matrix(f, with_methods=True)
    L_ref = oracle.fill(f)
    iu = np.triu_indices(len(f))
    _, m_ref = oracle.fill_pairs(f, iu[0], iu[1])
    same = L.view(np.uint64) == L_ref.view(np.uint64)
    assert same.all(), f"{(~same).sum()} entries differ, e.g. {np.argwhere(~same)[:3].tolist()}"
    assert np.array_equal(methods[iu].astype(np.int32), m_ref)
    assert np.array_equal(L, L.T)


@pytest.mark.parametrize("name", ["busbar", "spiral", "coupled", "meander15"])
def test_mesh_blocks_are_bit_identical(name):
    g, ports, d = _fixtures()[name]
    packed = model.pack(g, ports, d)
    system = MeshSystem.assemble_packed(packed, ports)
    ref = oracle.System(packed)
    got, want = system.blocks(), ref.blocks()
    for key in want:
        assert np.array_equal(got[key], want[key]), key
    system.close()


@pytest.mark.parametrize("name", ["busbar", "spiral", "coupled", "meander15", "cloud_manhattan"])
def test_fp64_impedance_matches_oracle(name):
    g, ports, d = _fixtures()[name]
    packed = model.pack(g, ports, d)
    freqs = np.concatenate([[0.0], fixtures.log_frequencies(1e3, 1e9, 11)])
    system = MeshSystem.assemble_packed(packed, ports, "fp64")
    z = system.sweep(freqs).impedance_ohm
    z_ref = oracle.System(packed).sweep(freqs)
    for k in range(len(freqs)):
        err = rel_frobenius(z[k], z_ref[k])
        assert err <= Z_TOL_FP64, f"{name} f={freqs[k]:.3g}: {err:.3e}"
    assert np.all(z[0].imag == 0.0)
    # reciprocity to rounding (tests/solve.rs:426-450)
    for k in range(len(freqs)):
        assert np.allclose(z[k], z[k].T, rtol=1e-12, atol=0)
    system.close()


def test_spiral_published_values():
    """docs/validation.md:107-112: L(1 kHz) = 0.637460 nH, R(DC) = 2.844828 ohm."""
    g, ports, d = fixtures.spiral_system()
    system = MeshSystem.assemble(g, ports, d)
    res = system.sweep([0.0, 1e3])
    assert round(res.resistance(0)[0, 0], 6) == 2.844828
    assert round(res.inductance(1)[0, 0] * 1e9, 6) == 0.637460
    system.close()


def test_invalid_frequency_reports_index():
    g, ports, d = fixtures.spiral_system()
    system = MeshSystem.assemble(g, ports, d)
    with pytest.raises(InvalidFrequency) as e:
        system.sweep([1e3, 2e3, -1.0, 5.0])
    assert e.value.index0 == 2
    with pytest.raises(InvalidFrequency):
        system.sweep([float("nan")])
    system.close()


def test_meander_2400_fp64_parity():
    """The 2 400-filament meander (benches/support.rs meander(24)), 21 log
    frequencies 10 kHz..1 GHz."""
    g, ports, d = fixtures.meander_problem(24)
    packed = model.pack(g, ports, d)
    freqs = fixtures.log_frequencies(1e4, 1e9, 21)
    system = MeshSystem.assemble_packed(packed, ports, "fp64")
    z = system.sweep(freqs).impedance_ohm
    ref = oracle.System(packed)
    assert np.array_equal(system.blocks()["l_ee"], ref.blocks()["l_ee"])
    z_ref = ref.sweep(freqs)
    errs = [rel_frobenius(z[k], z_ref[k]) for k in range(len(freqs))]
    assert max(errs) <= Z_TOL_FP64, errs
    system.close()


Z_TOL_REFINED = 1e-10


@pytest.mark.parametrize("name", ["busbar", "spiral", "coupled", "meander15", "cloud_manhattan"])
def test_refined_impedance_matches_oracle(name):
    """FP32 matrix-core factorization + FP64 refinement: Z within 1e-10."""
    g, ports, d = _fixtures()[name]
    packed = model.pack(g, ports, d)
    freqs = np.concatenate([[0.0], fixtures.log_frequencies(1e3, 1e9, 11)])
    system = MeshSystem.assemble_packed(packed, ports, "fp32_refined")
    res = system.sweep(freqs)
    z = res.impedance_ohm
    z_ref = oracle.System(packed).sweep(freqs)
    errs = [rel_frobenius(z[k], z_ref[k]) for k in range(len(freqs))]
    assert max(errs) <= Z_TOL_REFINED, errs
    assert np.all(z[0].imag == 0.0)
    if system.counts().internal_meshes > 0:
        assert res.extra["refine_steps_max"] >= 1
    system.close()


def test_meander_2400_refined_parity():
    g, ports, d = fixtures.meander_problem(24)
    packed = model.pack(g, ports, d)
    freqs = fixtures.log_frequencies(1e4, 1e9, 21)
    system = MeshSystem.assemble_packed(packed, ports, "fp32_refined")
    res = system.sweep(freqs)
    z = res.impedance_ohm
    z_ref = oracle.System(packed).sweep(freqs)
    errs = [rel_frobenius(z[k], z_ref[k]) for k in range(len(freqs))]
    assert max(errs) <= Z_TOL_REFINED, errs
    # the split-bf16 factorization (several 256-column outer blocks here)
    # was accurate enough that the FP32 fallback was not needed
    assert system.counts().internal_meshes > 512
    assert res.extra["factor_arith"] == 2
    assert res.extra["refine_steps_max"] <= 4
    system.close()


@pytest.mark.parametrize("parts", [1, 2, 3, 4])
def test_split_sweep_matches_one_frequency_at_a_time(parts):
    # a refined chunk on m >= 1024 is solved as up to `parts` groups of at
    # least 4 frequencies on as many streams; each frequency's Z is that of
    # its own sweep
    g, ports, d = fixtures.meander_problem(24)
    freqs = fixtures.log_frequencies(1e4, 1e9, 17)
    system = MeshSystem.assemble(g, ports, d, precision="fp32_refined")
    system.set_sweep_parts(parts)
    assert system.counts().internal_meshes >= 1024
    z = system.sweep(freqs).impedance_ohm
    for k in (0, 5, 8, 12, 16):
        z1 = system.sweep([freqs[k]]).impedance_ohm[0]
        assert rel_frobenius(z[k], z1) <= 1e-12, (k, freqs[k])
    system.set_precision("fp64")
    z64 = system.sweep(freqs).impedance_ohm
    assert max(rel_frobenius(z[k], z64[k]) for k in range(len(freqs))) <= Z_TOL_REFINED
    system.close()


def test_sweep_parts_argument():
    g, ports, d = fixtures.meander_problem(2)
    system = MeshSystem.assemble(g, ports, d, precision="fp32_refined")
    for bad in (0, 5, -1):
        with pytest.raises(InvalidArgument):
            system.set_sweep_parts(bad)
    system.close()
