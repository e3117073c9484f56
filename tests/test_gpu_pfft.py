"""GPU parity of the matrix-free path against the oracle, through the C ABI.

* pFFT operator (pfft/mod.rs:332-528): grid (spacing bit for bit, dims),
  projection CSR (indices equal, weights bit for bit), near-field
  corrections (pattern equal, values bit for bit: exact pair values from
  the pair fill, grid interactions in the reference's order of addition);
  y = L x within 1e-12 of the oracle (the FFT's rounding differs) and
  within the reference's 1e-6 of the dense GPU fill (tests/pfft.rs:44-95);
  symmetry and linearity (tests/pfft.rs:109-141); parameter errors.
* GMRES (gmres.rs:91-206) on the GPU: the reference's test systems
  (gmres.rs:259-411) against LAPACK and the oracle.
* Iterative system (iterative.rs): Z against the oracle's iterative system
  (both at GMRES tolerance 1e-10) and against the dense GPU solve < 1e-4
  (tests/solve.rs:775-832); DC real (:835-853); NotConverged reporting.
"""
import numpy as np
import pytest

import oracle
from fasterhenry_amd import (IterativeSystem, MeshSystem, NotConverged, PfftError,
                             PfftOperator, fixtures, gmres_dense, iterative_params, model,
                             partial_inductance_matrix, pfft_params)

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _portable():
    oracle.set_portable_math(True)
    yield


def _lcg_vector(n, seed):
    r = fixtures.lcg(seed)
    return np.array([r() - 0.5 for _ in range(n)])


def _random_set(n, manhattan, seed, side=None):
    g, ports, d = fixtures.random_cloud_problem(n, manhattan, seed, side)
    return oracle.discretize(model.pack(g, ports, d))


def _sets():
    sp = fixtures.spiral_system()
    return {
        "manhattan60": _random_set(60, True, 1),
        "skew60": _random_set(60, False, 1),
        "manhattan200": _random_set(200, True, 2),
        "skew200": _random_set(200, False, 2),
        "spiral": oracle.discretize(model.pack(*sp)),
    }


def _gpu_params(o):
    return pfft_params(spacing=None if o.auto_grid else o.spacing,
                       cells_per_filament=o.cells_per_filament,
                       near_field_radius=o.near_field_radius,
                       interpolation_order=o.interpolation_order)


@pytest.mark.parametrize("name", list(_sets().keys()))
def test_operator_data_is_bit_identical_to_the_oracle(name):
    f = _sets()[name]
    print(f"pfft data parity: {name}, n={len(f)}", flush=True)
    ref = oracle.Pfft(f)
    op = PfftOperator(f, _gpu_params(ref.params))
    want, got = ref.data(), op.data()
    assert got["grid_spacing"] == want["grid_spacing"]
    assert list(got["grid_dims"]) == list(want["grid_dims"])
    assert list(got["fft_dims"]) == list(want["fft_dims"])
    assert np.array_equal(got["proj_off"], want["proj_off"])
    assert np.array_equal(got["proj_idx"], want["proj_idx"])
    assert np.array_equal(got["proj_w"].view(np.uint64), want["proj_w"].view(np.uint64))
    assert np.array_equal(got["near_off"], want["near_off"])
    assert np.array_equal(got["near_col"], want["near_col"])
    bad = got["near_val"].view(np.uint64) != want["near_val"].view(np.uint64)
    assert not bad.any(), (np.flatnonzero(bad)[:5], got["near_val"][bad][:5], want["near_val"][bad][:5])


@pytest.mark.parametrize("name", list(_sets().keys()))
def test_apply_matches_the_oracle_and_the_dense_fill(name):
    f = _sets()[name]
    ref = oracle.Pfft(f)
    op = PfftOperator(f, _gpu_params(ref.params))
    x = _lcg_vector(len(f), 99)
    y = op.apply(x)
    yo = ref.apply(x)
    assert np.linalg.norm(y - yo) / np.linalg.norm(yo) < 1e-12
    dense = partial_inductance_matrix(f) @ x
    assert np.linalg.norm(y - dense) / np.linalg.norm(dense) < 1e-6
    assert np.max(np.abs(y - dense)) / np.max(np.abs(dense)) < 1e-6


@pytest.mark.parametrize("name", ["manhattan200", "skew200", "spiral"])
def test_fp32_convolution_stays_within_the_reference_gate(name):
    """FP32 far-field FFT: within the reference's 1e-6 of the dense fill and
    within 1e-6 of the FP64 operator (its near field and projection are the
    same FP64 data)."""
    f = _sets()[name]
    op64 = PfftOperator(f)
    op32 = PfftOperator(f, pfft_params(fft_precision="fp32"))
    assert op64.stats()["fft_kind"] == 0 and op32.stats()["fft_kind"] == 1
    d64, d32 = op64.data(), op32.data()
    assert np.array_equal(d64["near_val"], d32["near_val"])
    assert np.array_equal(d64["proj_w"], d32["proj_w"])
    x = _lcg_vector(len(f), 99)
    y64, y32 = op64.apply(x), op32.apply(x)
    assert np.linalg.norm(y32 - y64) / np.linalg.norm(y64) < 1e-6
    dense = partial_inductance_matrix(f) @ x
    assert np.linalg.norm(y32 - dense) / np.linalg.norm(dense) < 1e-6


def test_long_axis_uses_the_global_fft_and_matches_the_oracle():
    """A padded axis longer than one LDS line (two filaments 0.5 m apart on
    a 0.1 mm grid: M0 = 10000) takes the global-memory Stockham 3-D FFT."""
    g = model.Geometry()
    for x0 in (0.0, 0.5):
        a = g.add_node(model.Node(x0, 0.0, 0.0))
        b = g.add_node(model.Node(x0 + 1e-3, 0.0, 0.0))
        g.add_segment(model.SegmentDef(a, b, 2e-4, 1e-4, 5.8e7))
    f = oracle.discretize(model.pack(g, [], model.Discretization.uniform(1, 1)))
    ref = oracle.Pfft(f, oracle.PfftParams.default(spacing=1e-4))
    op = PfftOperator(f, pfft_params(spacing=1e-4))
    st = op.stats()
    assert st["fft_kind"] == 2 and st["fft_dims"][0] > 8191
    x = np.array([1.0, -0.5])
    y, yo = op.apply(x), ref.apply(x)
    assert np.linalg.norm(y - yo) / np.linalg.norm(yo) < 1e-12
    dense = partial_inductance_matrix(f) @ x
    assert np.linalg.norm(y - dense) / np.linalg.norm(dense) < 1e-6


def test_long_axis_cloud_matches_the_oracle_and_the_dense_fill():
    """Forty filaments strung along 0.43 m with offsets in y and z (every
    axis of the padded grid has several points; the long one is over 8 191),
    FP32 precision requested: the global FFT runs in FP64 and meets the
    oracle at 1e-12 and the dense fill at the reference's 1e-6."""
    rng = np.random.default_rng(8)
    g = model.Geometry()
    for k in range(40):
        x0 = 0.011 * k
        y0, z0 = rng.uniform(-1e-3, 1e-3, 2)
        a = g.add_node(model.Node(x0, y0, z0))
        b = g.add_node(model.Node(x0 + 2e-3, y0 + rng.uniform(-5e-4, 5e-4), z0))
        g.add_segment(model.SegmentDef(a, b, 1e-4, 5e-5, 5.8e7))
    f = oracle.discretize(model.pack(g, [], model.Discretization.uniform(1, 1)))
    ref = oracle.Pfft(f, oracle.PfftParams.default(spacing=1e-4))
    op = PfftOperator(f, pfft_params(spacing=1e-4, fft_precision="fp32"))
    st = op.stats()
    assert st["fft_kind"] == 2 and max(st["fft_dims"]) > 8191 and min(st["fft_dims"]) > 1
    x = _lcg_vector(len(f), 3)
    y, yo = op.apply(x), ref.apply(x)
    assert np.linalg.norm(y - yo) / np.linalg.norm(yo) < 1e-12
    dense = partial_inductance_matrix(f) @ x
    assert np.linalg.norm(y - dense) / np.linalg.norm(dense) < 1e-6
    op.close()


def test_batched_apply_equals_single_applies():
    f = _random_set(200, False, 4)
    op = PfftOperator(f)
    xs = np.stack([_lcg_vector(len(f), s) for s in (1, 2, 3)])
    ys = op.apply(xs)
    for k in range(3):
        y1 = op.apply(xs[k])
        assert np.linalg.norm(ys[k] - y1) <= 1e-13 * np.linalg.norm(y1)


def test_operator_is_symmetric_linear_and_repeatable():
    f = _random_set(300, False, 5, side=10e-3)
    op = PfftOperator(f)
    u, v = _lcg_vector(len(f), 1), _lcg_vector(len(f), 2)
    left, right = u @ op.apply(v), op.apply(u) @ v
    assert abs(left - right) <= 1e-12 * max(abs(left), abs(right))
    lu, lv, lc = op.apply(u), op.apply(v), op.apply(2.0 * u - 3.0 * v)
    assert np.max(np.abs(lc - (2.0 * lu - 3.0 * lv))) <= 1e-12 * np.max(np.abs(lc))
    assert np.array_equal(op.apply(u), lu)


def test_invalid_parameters_are_rejected():
    f = _random_set(10, True, 1)
    for params in (pfft_params(interpolation_order=0),
                   pfft_params(near_field_radius=3.0, interpolation_order=5),
                   pfft_params(cells_per_filament=0.0),
                   pfft_params(spacing=-1.0)):
        with pytest.raises(PfftError) as e:
            PfftOperator(f, params)
        assert e.value.code == 1
    with pytest.raises(PfftError) as e:
        PfftOperator(f, pfft_params(spacing=1e-7))
    assert e.value.code == 2 and e.value.index1 == 1 << 26


def test_empty_operator():
    op = PfftOperator(np.zeros(0, dtype=model.FILAMENT_DTYPE))
    assert op.stats()["filaments"] == 0


# ---------------------------------------------------------------------------
# GMRES on the device

def _stream(seed):
    mask = (1 << 64) - 1
    state = ((seed * 0x9E3779B97F4A7C15) & mask) | 1

    def nxt():
        nonlocal state
        state ^= (state << 13) & mask
        state ^= state >> 7
        state ^= (state << 17) & mask
        return (state >> 11) / float(1 << 52) - 1.0
    return nxt


def _complex_symmetric(n, seed):
    nxt = _stream(seed)
    a = np.zeros((n, n), dtype=np.complex128)
    for j in range(n):
        for i in range(n):
            re = nxt()
            a[i, j] = complex(re, nxt()) * 0.3
    a = a + a.T
    a[np.diag_indices(n)] += complex(n, 0.5 * n)
    return a


def test_gmres_complex_system_matches_lapack_and_the_oracle():
    n = 60
    a = _complex_symmetric(n, 3)
    nxt = _stream(7)
    b = np.array([complex(nxt(), nxt()) for _ in range(n)])
    x, its, rr, conv = gmres_dense(a, b, restart=8)
    xo, itso, rro, convo = oracle.gmres_dense(a, b, restart=8)
    assert conv and rr <= 1e-10 and its == itso
    direct = np.linalg.solve(a, b)
    assert np.linalg.norm(x - direct) / np.linalg.norm(direct) < 1e-9
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-12


def test_gmres_real_system_stays_real():
    n = 40
    nxt = _stream(11)
    a = np.zeros((n, n))
    for j in range(n):
        for i in range(n):
            a[i, j] = nxt()
    a[np.diag_indices(n)] += 2.0 * n
    b = np.array([nxt() for _ in range(n)])
    x, its, rr, conv = gmres_dense(a, b, complex_arith=False)
    assert conv and np.all(x.imag == 0.0)
    direct = np.linalg.solve(a, b)
    assert np.linalg.norm(x.real - direct) < 1e-9 * np.linalg.norm(direct)


def test_gmres_full_krylov_space_zero_rhs_and_iteration_cap():
    n = 12
    a = _complex_symmetric(n, 5)
    b = np.array([complex(i, 1.0) for i in range(n)])
    x, its, rr, conv = gmres_dense(a, b, tolerance=1e-13, restart=n, max_iterations=n)
    assert conv and its <= n
    x, its, rr, conv = gmres_dense(np.eye(4, dtype=complex), np.zeros(4, dtype=complex),
                                   x0=[3 + 1j] * 4)
    assert its == 0 and conv and np.all(x == 0)
    a = _complex_symmetric(30, 13)
    x, its, rr, conv = gmres_dense(a, np.ones(30, dtype=complex), tolerance=1e-15, restart=2,
                                   max_iterations=3)
    assert not conv and its == 3 and rr < 1.0


# ---------------------------------------------------------------------------
# the iterative system

FREQS = [0.0, 1e3, 1e6, 1e9, 1e10]


def test_iterative_system_matches_the_oracle_and_the_dense_solve():
    g, ports, d = fixtures.coupled_structure()
    packed = model.pack(g, ports, d)
    it = IterativeSystem.assemble_packed(packed, ports)
    z, info = it.sweep_raw(FREQS)
    ref = oracle.IterSystem(packed)
    dense = MeshSystem.assemble_packed(packed, ports, "fp64")
    zd, _ = dense.sweep_raw(FREQS)
    assert info.solver == 1 and info.gmres_iterations > 0
    for k, f in enumerate(FREQS):
        zo, its = ref.impedance(f)
        assert np.max(np.abs(z[k] - zo)) / np.max(np.abs(zo)) < 1e-8, f
        assert np.max(np.abs(z[k] - zd[k])) / np.max(np.abs(zd[k])) < 1e-4, f
    assert np.all(z[0].imag == 0.0)
    assert z[0][0, 1] == 0 and z[0][0, 2] == 0


def test_iterative_meander_matches_dense():
    g, ports, d = fixtures.meander_problem(15)
    packed = model.pack(g, ports, d)
    it = IterativeSystem.assemble_packed(packed, ports, iterative_params(tolerance=1e-8, restart=30))
    dense = MeshSystem.assemble_packed(packed, ports, "fp64")
    freqs = [1e6, 1e8, 1e9]
    z, info = it.sweep_raw(freqs)
    zd, _ = dense.sweep_raw(freqs)
    for k in range(len(freqs)):
        assert np.linalg.norm(z[k] - zd[k]) / np.linalg.norm(zd[k]) < 1e-4
    assert info.gmres_iterations_max < 50


def test_gmres_not_converged_reports_port_and_frequency():
    g, ports, d = fixtures.coupled_structure()
    packed = model.pack(g, ports, d)
    it = IterativeSystem.assemble_packed(packed, ports,
                                         iterative_params(tolerance=1e-15, max_iterations=1))
    with pytest.raises(NotConverged) as e:
        it.sweep_raw([1e3, 1e6])
    assert e.value.index0 == 0 and e.value.index1 == 0 and e.value.code == 1
    assert 0.0 < e.value.value < 1.0
