"""GPU parity at the full sizes of the measured workloads.

* Dense sweep, meander(49) at 2 x 2 (9 800 filaments, 7 350 internal loops,
  41 log frequencies 10 kHz - 1 GHz; benches/scaling.rs:91): sampled rows
  of L bit for bit against the oracle's pair fill; the FP64 GPU Z within
  1e-12 of the oracle's LU + Schur (dense.rs:43-189, solve.rs:1154-1164) at
  two frequencies; the refined form within 1e-10 of the FP64 form at every
  frequency; the matrix-free path within 1e-4 of the dense one.
* Batched sweep: all 48 seeded geometries x 16 frequencies (100 kHz -
  3 GHz), each bit-identical to its own solve; three surface-graded members
  against the oracle within 1e-12.
* Adjoint: the 6-turn spiral (about 6 000 filaments, 12 parameters, 5
  frequencies) against central differences of the device's Z at relative
  step 1e-6, within 1e-6 of the largest derivative per frequency.
* Matrix-free: meander(157) (99 224 filaments, grid 899 x 898 x 8, padded
  1800 x 1800 x 15): the FP32-FFT product within 1e-6 of the FP64-FFT one;
  64 sampled rows within 1e-6 of the exact rows sum_j L_ij x_j from the
  oracle's pair fill (the reference's gate, tests/pfft.rs:44-95); symmetry
  and linearity (tests/pfft.rs:109-141).
"""
import numpy as np
import pytest

import oracle
from fasterhenry_amd import (IterativeSystem, MeshSystem, PfftOperator, fixtures, iterative_params,
                             model, pfft_params, solve_batch)

pytestmark = pytest.mark.gpu

HEADLINE_FREQS = fixtures.log_frequencies(1e4, 1e9, 41)


def rel_frobenius(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(autouse=True)
def _portable():
    oracle.set_portable_math(True)
    yield


@pytest.fixture(scope="module")
def headline():
    g, ports, d = fixtures.meander_problem(49)
    packed = model.pack(g, ports, d)
    print(f"headline: assembling meander(49), {packed.filament_count} filaments", flush=True)
    system = MeshSystem.assemble_packed(packed, ports, "fp64")
    yield packed, ports, system
    system.close()


@pytest.mark.timeout(600)
def test_headline_fill_rows_are_bit_identical(headline):
    packed, _, system = headline
    f = system.filaments()
    assert len(f) == 9800 and system.counts().internal_meshes == 7350
    assert f.tobytes() == oracle.discretize(packed).tobytes()
    L = system.inductance_matrix()
    rows = np.concatenate([np.arange(0, 9800, 197), [1, 2, 3, 9797, 9798, 9799]])
    r = np.repeat(rows, 9800).astype(np.int64)
    c = np.tile(np.arange(9800), len(rows)).astype(np.int64)
    want, _ = oracle.fill_pairs(f, r, c)
    got = L[rows].ravel()
    same = got.view(np.uint64) == want.view(np.uint64)
    assert same.all(), f"{(~same).sum()} of {len(same)} sampled entries differ"
    assert np.array_equal(L, L.T)


@pytest.mark.timeout(600)
def test_headline_fp64_matches_the_oracle_lu(headline):
    """Two frequencies through the oracle's LU (PANEL 32, |re|+|im| pivots)
    and Schur complement on the mesh blocks."""
    _, _, system = headline
    z = system.sweep(HEADLINE_FREQS).impedance_ohm
    b = system.blocks()
    for k in (0, 30):
        w = 2 * np.pi * HEADLINE_FREQS[k]
        ee = b["r_ee"] + 1j * w * b["l_ee"]
        ep = b["r_ep"] + 1j * w * b["l_ep"]
        pp = b["r_pp"] + 1j * w * b["l_pp"]
        print(f"headline: oracle LU at {HEADLINE_FREQS[k]:.4g} Hz", flush=True)
        x = oracle.lu_solve(ee, ep)
        zr = pp - ep.T @ x
        err = rel_frobenius(z[k], zr)
        print(f"headline: fp64 vs oracle {err:.2e}", flush=True)
        assert err <= 1e-12


@pytest.mark.timeout(600)
def test_headline_refined_matches_fp64_at_every_frequency(headline):
    _, _, system = headline
    z64 = system.sweep(HEADLINE_FREQS).impedance_ohm
    system.set_precision("fp32_refined")
    try:
        res = system.sweep(HEADLINE_FREQS)
    finally:
        system.set_precision("fp64")
    err = [rel_frobenius(res.impedance_ohm[k], z64[k]) for k in range(len(HEADLINE_FREQS))]
    print(f"headline: refined vs fp64 max {max(err):.2e}", flush=True)
    assert max(err) <= 1e-10
    assert np.all(res.impedance_ohm[:, 0, 0].real > 0)


@pytest.mark.timeout(600)
def test_headline_matrix_free_agrees_with_dense(headline):
    """The iterative path (FP32 FFT, GMRES restart 30, tolerance 1e-8) on the
    same 9 800 filaments within 1e-4 of the dense solve (tests/solve.rs:775-832)."""
    packed, ports, system = headline
    freqs = fixtures.log_frequencies(1e6, 1e9, 4)
    zd = system.sweep(freqs).impedance_ohm
    params = iterative_params(pfft_params(fft_precision="fp32"), tolerance=1e-8, restart=30)
    it = IterativeSystem.assemble_packed(packed, ports, params)
    zi = it.sweep(freqs).impedance_ohm
    it.close()
    err = max(rel_frobenius(zi[k], zd[k]) for k in range(len(freqs)))
    print(f"headline: iterative vs dense {err:.2e}", flush=True)
    assert err < 1e-4


# --------------------------------------------------------------------------
# The 2 400-filament scaling fixture and its trace variant

C2_FREQS = fixtures.log_frequencies(1e4, 1e9, 21)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("name", ["meander24", "trace_meander"])
def test_c2_sweeps_match_the_oracle(name):
    """meander(24) at 2 x 2 (2 400 filaments, the README's scaling fixture)
    and the 0.2 mm x 35 um trace of 12 runs of 8 mm at 0.4 mm pitch, graded
    5 x 3 (2 340 filaments): FP64 GPU within 1e-12 of the oracle's LU + Schur
    at five frequencies, the refined form within 1e-10 of FP64 at all 21."""
    prob = fixtures.meander_problem(24) if name == "meander24" else fixtures.trace_meander()
    packed = model.pack(*prob)
    assert packed.filament_count == (2400 if name == "meander24" else 2340)
    system = MeshSystem.assemble_packed(packed, prob[1], "fp64")
    z64 = system.sweep(C2_FREQS).impedance_ohm
    system.set_precision("fp32_refined")
    zr = system.sweep(C2_FREQS).impedance_ohm
    system.close()
    sub = [0, 5, 10, 15, 20]
    zo = oracle.System(packed).sweep(C2_FREQS[sub])
    for j, k in enumerate(sub):
        assert rel_frobenius(z64[k], zo[j]) <= 1e-12, (name, k)
    err = max(rel_frobenius(zr[k], z64[k]) for k in range(len(C2_FREQS)))
    print(f"{name}: refined vs fp64 {err:.2e}", flush=True)
    assert err <= 1e-10


# --------------------------------------------------------------------------
# The batched sweep at full size

BATCH_FREQS = fixtures.log_frequencies(1e5, 3e9, 16)


@pytest.mark.timeout(900)
def test_full_batch_equals_one_solve_per_problem():
    probs = fixtures.batch_problems()
    counts = [model.pack(*p).filament_count for p in probs]
    assert len(probs) == 48 and min(counts) >= 300 and max(counts) <= 4000
    print(f"batch: 48 problems, {sum(counts)} filaments", flush=True)
    zs, infos = solve_batch(probs, BATCH_FREQS, precision="fp32_refined", concurrency=4)
    for k, ((g, ports, d), z, info) in enumerate(zip(probs, zs, infos)):
        packed = model.pack(g, ports, d)
        system = MeshSystem.assemble_packed(packed, ports, "fp32_refined")
        z1, _ = system.sweep_raw(BATCH_FREQS)
        system.close()
        assert np.array_equal(z.view(np.float64), z1.view(np.float64)), k
        assert info["filaments"] == counts[k]
    print("batch: every problem bit-identical to its own solve", flush=True)


@pytest.mark.timeout(300)
def test_full_fp64_batch_on_four_workers_equals_one_worker():
    # Regression: the panel kernel's in-place stores once overtook a
    # sibling wave's loads of the same rows when the waves drifted apart
    # under the other workers' kernels (a wrong 32 x 32 block in about one
    # FP64 batch in four).  Several batches, each bit-identical.
    probs = fixtures.batch_problems()
    ref, _ = solve_batch(probs, BATCH_FREQS, precision="fp64", concurrency=1)
    for rep in range(4):
        zs, _ = solve_batch(probs, BATCH_FREQS, precision="fp64", concurrency=4)
        for k, (z, z1) in enumerate(zip(zs, ref)):
            assert np.array_equal(z.view(np.float64), z1.view(np.float64)), (rep, k)
    print("batch: four FP64 batches on four workers bit-identical to one worker", flush=True)


@pytest.mark.timeout(900)
def test_full_batch_graded_members_match_the_oracle():
    probs = fixtures.batch_problems()
    counts = [model.pack(*p).filament_count for p in probs]
    order = np.argsort(counts)
    pick = [int(order[0]), int(order[1]), int(order[len(order) // 3])]
    zs64, _ = solve_batch(probs, BATCH_FREQS, precision="fp64", concurrency=4)
    zsr, infos = solve_batch(probs, BATCH_FREQS, precision="fp32_refined", concurrency=4)
    for k in range(len(probs)):
        err = max(rel_frobenius(zsr[k][j], zs64[k][j]) for j in range(len(BATCH_FREQS)))
        if err > 1e-10:
            # name the side that moved: each against its own single solve
            packed = model.pack(*probs[k])
            system = MeshSystem.assemble_packed(packed, probs[k][1], "fp64")
            z1, _ = system.sweep_raw(BATCH_FREQS)
            system.set_precision("fp32_refined")
            zr1, _ = system.sweep_raw(BATCH_FREQS)
            system.close()
            e64 = max(rel_frobenius(zs64[k][j], z1[j]) for j in range(len(BATCH_FREQS)))
            er = max(rel_frobenius(zsr[k][j], zr1[j]) for j in range(len(BATCH_FREQS)))
            print(f"problem {k}: batch fp64 vs single {e64:.2e}, batch refined vs single {er:.2e}, "
                  f"info {infos[k]}", flush=True)
        assert err <= 1e-10, (k, err)
    sub = [0, 5, 10, 15]
    for k in pick:
        packed = model.pack(*probs[k])
        assert packed.grids["graded"].all()
        print(f"batch: oracle solve of problem {k} ({counts[k]} filaments)", flush=True)
        zr = oracle.System(packed).sweep(BATCH_FREQS[sub])
        for j, kk in enumerate(sub):
            assert rel_frobenius(zs64[k][kk], zr[j]) <= 1e-12, (k, kk)


# --------------------------------------------------------------------------
# Adjoint gradients at full size

ADJOINT_FREQS = [1e6, 1e7, 1e8, 5e8, 1e9]


def _z_tight(system, freqs):
    nn = len(system._packed.nodes)
    return system.grad(freqs, np.zeros((0, nn, 3)))[0]


@pytest.mark.timeout(900)
def test_full_adjoint_gradient_matches_device_central_differences():
    s = fixtures.AdjointSpiral()
    g, ports, d = s.problem()
    system = MeshSystem.assemble(g, ports, d, precision="fp32_refined")
    n = system.counts().filaments
    assert 5500 <= n <= 6500 and s.n_params == 12
    print(f"adjoint: {n} filaments, 12 parameters", flush=True)
    d_nodes, d_widths = s.directions()
    z, dr, dl = system.grad(ADJOINT_FREQS, d_nodes, d_widths)
    system.close()
    w = 2 * np.pi * np.asarray(ADJOINT_FREQS)
    fr = np.zeros((len(ADJOINT_FREQS), s.n_params))
    fl = np.zeros_like(fr)
    for q in range(s.n_params):
        h = 1e-6 * s.nominal[q]
        zz = []
        for sgn in (1.0, -1.0):
            p = s.nominal.copy()
            p[q] += sgn * h
            sysq = MeshSystem.assemble(*s.problem(p), precision="fp32_refined")
            zz.append(_z_tight(sysq, ADJOINT_FREQS)[:, 0, 0])
            sysq.close()
        dz = (zz[0] - zz[1]) / (2 * h)
        fr[:, q], fl[:, q] = dz.real, dz.imag / w
    er = np.abs(dr[:, :, 0, 0] - fr).max(axis=1) / np.abs(fr).max(axis=1)
    el = np.abs(dl[:, :, 0, 0] - fl).max(axis=1) / np.abs(fl).max(axis=1)
    print("adjoint: dR err", er, "dL err", el, flush=True)
    assert np.all(er < 1e-6) and np.all(el < 1e-6)


# --------------------------------------------------------------------------
# The matrix-free operator at full size

@pytest.fixture(scope="module")
def meander157():
    g, ports, d = fixtures.meander_problem(157)
    packed = model.pack(g, ports, d)
    f = oracle.discretize(packed)
    assert len(f) == 99224
    print("pfft: building the FP32 and FP64 operators of meander(157)", flush=True)
    op32 = PfftOperator(f, pfft_params(fft_precision="fp32"))
    op64 = PfftOperator(f, pfft_params(fft_precision="fp64"))
    yield f, op32, op64
    op32.close()
    op64.close()


def _lcg_vector(n, seed):
    r = fixtures.lcg(seed)
    return np.array([r() - 0.5 for _ in range(n)])


@pytest.mark.timeout(600)
def test_full_pfft_grid_and_fft_precisions(meander157):
    f, op32, op64 = meander157
    st = op32.stats()
    assert tuple(st["fft_dims"]) == (1800, 1800, 15)
    assert st["fft_kind"] == 1 and op64.stats()["fft_kind"] == 0
    x = _lcg_vector(len(f), 99)
    y32, y64 = op32.apply(x), op64.apply(x)
    err = np.linalg.norm(y32 - y64) / np.linalg.norm(y64)
    print(f"pfft: fp32 vs fp64 FFT {err:.2e}", flush=True)
    assert err <= 1e-6


@pytest.mark.timeout(600)
def test_full_pfft_sampled_rows_match_the_exact_product(meander157):
    f, op32, op64 = meander157
    n = len(f)
    x = _lcg_vector(n, 99)
    rng = np.random.default_rng(157)
    rows = np.unique(np.concatenate([np.linspace(0, n - 1, 32).astype(np.int64),
                                     rng.integers(0, n, 32)]))
    r = np.repeat(rows, n).astype(np.int64)
    c = np.tile(np.arange(n), len(rows)).astype(np.int64)
    vals, _ = oracle.fill_pairs(f, r, c)
    exact = vals.reshape(len(rows), n) @ x
    for name, op in (("fp32", op32), ("fp64", op64)):
        y = op.apply(x)[rows]
        norm = np.linalg.norm(y - exact) / np.linalg.norm(exact)
        amax = np.max(np.abs(y - exact)) / np.max(np.abs(exact))
        print(f"pfft {name}: sampled rows {norm:.2e} (norm) {amax:.2e} (max)", flush=True)
        assert norm < 1e-6 and amax < 1e-6


@pytest.mark.timeout(600)
def test_full_pfft_symmetric_and_linear(meander157):
    f, op32, op64 = meander157
    u, v = _lcg_vector(len(f), 1), _lcg_vector(len(f), 2)
    for op, tol in ((op64, 1e-12), (op32, 1e-6)):
        lu, lv = op.apply(u), op.apply(v)
        # u . Lv = Lu . v, on the scale of the vectors (the inner products of
        # sign-alternating vectors cancel)
        left, right = u @ lv, lu @ v
        assert abs(left - right) <= tol * np.linalg.norm(u) * np.linalg.norm(lv)
        lc = op.apply(2.0 * u - 3.0 * v)
        assert np.max(np.abs(lc - (2.0 * lu - 3.0 * lv))) <= tol * np.max(np.abs(lc))
        assert np.array_equal(op.apply(u), lu)
