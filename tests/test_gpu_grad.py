"""Gradients on the device (fh_grad, fh_partial_inductance_derivative).

* Pair level: the nominal values of the derivative kernels are the fill's
  bits, and their derivatives match central differences of the oracle's
  mutual inductance along the filament tangents, for every method of the
  dispatch (orthogonal, point quadrature inline and on a wave, closed form,
  aligned quadrature, sampled filaments).
* Solve level: dR/dp and dL/dp of the port impedance match central
  differences (relative step 1e-6) of the oracle's solve on a small spiral
  and of the device's FP64 solve on a larger one, within 1e-6 of the largest
  derivative of each frequency.
"""
import numpy as np
import pytest

import oracle
from fasterhenry_amd import (InvalidArgument, MeshSystem, SolveError, api, fixtures, model,
                             partial_inductance_derivative, partial_inductance_matrix)

pytestmark = pytest.mark.gpu

FREQS = [1e6, 1e7, 1e8, 5e8, 1e9]
FIELDS = ["start", "end", "length", "dir", "wdir", "hdir", "width", "height"]


def _shifted(f, t, h):
    out = f.copy()
    for k in FIELDS:
        out[k] = np.asarray(f[k]) + h * np.asarray(t[k])
    return out


def _tangents(spiral, packed):
    d_nodes, d_widths = spiral.directions()
    return np.stack([api.discretize_tangent(packed, d_nodes[q], d_widths[q])
                     for q in range(spiral.n_params)])


def _pick(methods, per, rng):
    n = methods.shape[0]
    iu = np.triu_indices(n)
    out = []
    for code in range(5):
        sel = np.nonzero(methods[iu] == code)[0]
        if len(sel):
            for s in rng.choice(sel, size=min(per, len(sel)), replace=False):
                out.append((iu[0][s], iu[1][s]))
    return out


def _check_pairs(F, T, pairs):
    L, methods = partial_inductance_matrix(F, with_methods=True)
    rows = np.array([(f, g, q) for (f, g, q) in pairs], dtype=np.uint32)
    v, d = partial_inductance_derivative(F, T, rows)
    seen = set()
    for k, (f, g, q) in enumerate(pairs):
        seen.add(int(methods[f, g]))
        assert v[k] == L[f, g], (f, g)  # the fill's bits
        # Central differences at steps of 1e-5, 1e-4 and 1e-3 of the pair's
        # smallest extent over the tangent's size: a pair near a switch of
        # method or order (closed form / quadrature, both accurate to about
        # 1e-11) spoils the smallest step, rounding and curvature the others;
        # a wrong derivative misses all three.
        size = min(F[f]["width"], F[f]["height"], F[f]["length"], F[g]["width"], F[g]["height"],
                   F[g]["length"])
        speed = max(np.abs(np.asarray(T[q][f][k2])).max() for k2 in FIELDS) + \
            max(np.abs(np.asarray(T[q][g][k2])).max() for k2 in FIELDS)
        if speed == 0.0:
            assert d[k] == 0.0  # neither filament moves
            continue
        errs = []
        for rel in (1e-5, 1e-4, 1e-3):
            h = rel * size / speed
            try:
                lp = oracle.mutual(_shifted(F[f], T[q][f], h), _shifted(F[g], T[q][g], h))[0]
                lm = oracle.mutual(_shifted(F[f], T[q][f], -h), _shifted(F[g], T[q][g], -h))[0]
            except oracle.OracleError:
                continue  # a shifted pair the oracle refuses (coincident sample lines)
            fd = (lp - lm) / (2 * h)
            # relative to the derivative, with a floor at 1e-6 of |L| / size
            scale = max(abs(fd), abs(v[k]) * speed / size)
            errs.append(abs(d[k] - fd) / scale if scale > 0.0 else abs(d[k]))
        assert errs and min(errs) <= 1e-6, (f, g, q, int(methods[f, g]), d[k], errs)
    return seen


def test_pair_derivatives_on_the_spiral():
    oracle.set_portable_math(True)
    s = fixtures.AdjointSpiral(turns=2, filaments=500)
    packed = model.pack(*s.problem())
    F = api.discretize_packed(packed)
    T = _tangents(s, packed)
    _, methods = partial_inductance_matrix(F, with_methods=True)
    rng = np.random.default_rng(5)
    pairs = [(f, g, int(rng.integers(s.n_params))) for (f, g) in _pick(methods, 12, rng)]
    print(f"spiral pairs: {len(pairs)}", flush=True)
    seen = _check_pairs(F, T, pairs)
    assert {0, 2, 3} <= seen


def _random_tangents(F, rng, count, turning):
    """Random directions: each filament moved, stretched along its axis and
    resized; with ``turning`` its frame vectors turn as well."""
    n = len(F)
    T = np.zeros((count, n), dtype=F.dtype)
    axis = np.asarray(F["dir"])
    for q in range(count):
        u = rng.standard_normal((n, 3))
        a, b = rng.standard_normal(n), rng.standard_normal(n)
        T[q]["start"] = u + a[:, None] * axis
        T[q]["end"] = u + b[:, None] * axis
        T[q]["length"] = b - a
        T[q]["width"] = 0.3 * rng.standard_normal(n)
        T[q]["height"] = 0.3 * rng.standard_normal(n)
        if turning:
            for k in ("dir", "wdir", "hdir"):
                T[q][k] = rng.standard_normal((n, 3))
    return T


@pytest.mark.parametrize("case", ["skew_bars", "busbars"])
def test_pair_derivatives_with_random_tangents(case):
    # skew bars (turning frames): sampled filaments and skew point
    # quadrature; busbars (moving and resizing only): the aligned closed form
    oracle.set_portable_math(True)
    rng = np.random.default_rng(3)
    if case == "skew_bars":
        g, ports, d = fixtures.random_cloud_problem(60, manhattan=False, seed=11, side=1.5e-3)
        F = api.discretize_packed(model.pack(g, ports, d))
        T = _random_tangents(F, rng, 2, turning=True)
        expect = 4
    else:
        g, ports, d = fixtures.busbar_pair(segments_per_bar=3, nw=4, nh=2)
        F = api.discretize_packed(model.pack(g, ports, d))
        T = _random_tangents(F, rng, 2, turning=False)
        expect = 1
    _, methods = partial_inductance_matrix(F, with_methods=True)
    pairs = [(f, g, int(rng.integers(2))) for (f, g) in _pick(methods, 10, rng)]
    if case == "skew_bars":
        # turning frames: only the methods that hold for any orientation
        pairs = [(f, g, q) for (f, g, q) in pairs if methods[f, g] in (3, 4)]
    print(f"{case} pairs: {len(pairs)}", flush=True)
    seen = _check_pairs(F, T, pairs)
    assert expect in seen


def test_turning_a_perpendicular_or_parallel_pair_is_refused():
    g, ports, d = fixtures.busbar_pair(segments_per_bar=2, nw=2, nh=1)
    F = api.discretize_packed(model.pack(g, ports, d))
    T = _random_tangents(F, np.random.default_rng(1), 1, turning=True)
    _, methods = partial_inductance_matrix(F, with_methods=True)
    f, gg = np.argwhere(methods == 1)[0]
    with pytest.raises(SolveError):
        partial_inductance_derivative(F, T, [(f, gg, 0)])
    # the whole system: a node moved across its segment turns it
    s = fixtures.AdjointSpiral(turns=2, filaments=200)
    system = MeshSystem.assemble(*s.problem(), precision="fp32_refined")
    d_nodes, _ = s.directions()
    turn = np.zeros_like(d_nodes[:1])
    turn[0, 1, 1] = 1.0  # node 1 lies on the first (x-directed) side: move it in y
    with pytest.raises(SolveError):
        system.grad(FREQS[:1], turn)
    system.close()


def _rel_err(g, fd):
    """max over parameters of |g - fd| / max |fd|, per frequency."""
    return np.abs(g - fd).max(axis=1) / np.abs(fd).max(axis=1)


def _fd_oracle(spiral, freqs, rel=1e-4):
    """Central differences of the oracle's Z, Richardson-extrapolated from
    relative steps rel and 2 rel: the oracle's LU leaves about 1e-11 of
    rounding in Z, which a step of 1e-6 would magnify to 1e-5."""
    P = spiral.n_params
    w = 2 * np.pi * np.asarray(freqs)
    out = []
    for step in (rel, 2 * rel):
        dz = np.zeros((len(freqs), P), dtype=np.complex128)
        for q in range(P):
            h = step * spiral.nominal[q]
            z = []
            for sgn in (1.0, -1.0):
                p = spiral.nominal.copy()
                p[q] += sgn * h
                z.append(oracle.System(model.pack(*spiral.problem(p))).sweep(freqs)[:, 0, 0])
            dz[:, q] = (z[0] - z[1]) / (2 * h)
        out.append(dz)
    dz = (4.0 * out[0] - out[1]) / 3.0
    return dz.real, dz.imag / w[:, None]


def test_gradient_matches_oracle_finite_differences():
    oracle.set_portable_math(True)
    s = fixtures.AdjointSpiral(turns=3, filaments=700)
    g, ports, d = s.problem()
    system = MeshSystem.assemble(g, ports, d, precision="fp32_refined")
    d_nodes, d_widths = s.directions()
    print(f"gradient: {system.counts().filaments} filaments, {s.n_params} params", flush=True)
    z, dr, dl = system.grad(FREQS, d_nodes, d_widths)
    system.close()
    zr = oracle.System(model.pack(g, ports, d)).sweep(FREQS)
    assert np.abs(z - zr).max() / np.abs(zr).max() < 1e-10
    fr, fl = _fd_oracle(s, FREQS)
    er, el = _rel_err(dr[:, :, 0, 0], fr), _rel_err(dl[:, :, 0, 0], fl)
    print("dR err", er, "dL err", el, flush=True)
    assert np.all(er < 1e-6) and np.all(el < 1e-6)


def _z_tight(system, freqs):
    """Z with x refined to working accuracy (fh_grad with no directions)."""
    nn = len(system._packed.nodes)
    return system.grad(freqs, np.zeros((0, nn, 3)))[0]


def test_gradient_matches_device_finite_differences():
    # relative step 1e-6 on the device's own Z, refined to working accuracy
    s = fixtures.AdjointSpiral(turns=4, filaments=2500)
    g, ports, d = s.problem()
    system = MeshSystem.assemble(g, ports, d, precision="fp32_refined")
    d_nodes, d_widths = s.directions()
    z, dr, dl = system.grad(FREQS, d_nodes, d_widths)
    z2, dr2, dl2 = system.grad(FREQS, d_nodes, d_widths)
    system.close()
    # deterministic
    assert np.array_equal(dr, dr2) and np.array_equal(dl, dl2) and np.array_equal(z, z2)
    w = 2 * np.pi * np.asarray(FREQS)
    fr = np.zeros((len(FREQS), s.n_params))
    fl = np.zeros_like(fr)
    for q in range(s.n_params):
        h = 1e-6 * s.nominal[q]
        zz = []
        for sgn in (1.0, -1.0):
            p = s.nominal.copy()
            p[q] += sgn * h
            sysq = MeshSystem.assemble(*s.problem(p), precision="fp32_refined")
            zz.append(_z_tight(sysq, FREQS)[:, 0, 0])
            sysq.close()
        dz = (zz[0] - zz[1]) / (2 * h)
        fr[:, q], fl[:, q] = dz.real, dz.imag / w
    er, el = _rel_err(dr[:, :, 0, 0], fr), _rel_err(dl[:, :, 0, 0], fl)
    print("dR err", er, "dL err", el, flush=True)
    assert np.all(er < 1e-6) and np.all(el < 1e-6)


def test_dc_resistance_gradient_has_a_closed_form():
    # at f = 0 the spiral is its segments in series, each l / (sigma w t)
    s = fixtures.AdjointSpiral(turns=3, filaments=700)
    system = MeshSystem.assemble(*s.problem(), precision="fp32_refined")
    d_nodes, d_widths = s.directions()
    z, dr, _ = system.grad([0.0], d_nodes, d_widths)
    system.close()

    def r_dc(params, grad=False):
        xyz = s.node_positions(params)
        w = np.array([params[k] for k in s.segment_turns()])
        seg = np.linalg.norm(xyz[1:] - xyz[:-1], axis=1)
        return (seg / (fixtures.COPPER * w * s.thickness)).sum()

    assert abs(z[0, 0, 0].real - r_dc(s.nominal)) <= 1e-12 * r_dc(s.nominal)
    # node positions are affine in p: the derivative of each segment length
    # follows from the node velocities, the width term from w
    xyz = s.node_positions(s.nominal)
    turn = s.segment_turns()
    w = np.array([s.nominal[k] for k in turn])
    seg = xyz[1:] - xyz[:-1]
    length = np.linalg.norm(seg, axis=1)
    for q in range(s.n_params):
        dseg = d_nodes[q][1:] - d_nodes[q][:-1]
        dlen = (seg * dseg).sum(axis=1) / length
        exact = ((dlen / w - length * d_widths[q] / w ** 2) / (fixtures.COPPER * s.thickness)).sum()
        assert abs(dr[0, q, 0, 0] - exact) <= 1e-9 * np.abs(dr[0]).max(), q


def test_gradient_arguments():
    s = fixtures.AdjointSpiral(turns=2, filaments=200)
    g, ports, d = s.problem()
    system = MeshSystem.assemble(g, ports, d, precision="fp32_refined")
    d_nodes, d_widths = s.directions()
    with pytest.raises(InvalidArgument):
        system.grad(FREQS, d_nodes[:, :-1], d_widths)
    with pytest.raises(InvalidArgument):
        system.grad(FREQS, d_nodes, d_widths[:, :-1])
    # DC: dL is 0 by definition
    z, dr, dl = system.grad([0.0, 1e6], d_nodes, d_widths)
    assert np.all(dl[0] == 0.0) and np.all(np.isfinite(dr))
    assert np.abs(dl[1]).max() > 0.0
    # a rigid translation of the whole spiral changes nothing
    shift = np.zeros((1,) + d_nodes.shape[1:])
    shift[0, :, 0] = 1.0
    _, dr_t, dl_t = system.grad([1e6], shift)
    assert np.abs(dr_t).max() <= 1e-9 * np.abs(dr[1]).max()
    assert np.abs(dl_t).max() <= 1e-9 * np.abs(dl[1]).max()
    system.close()
    it = MeshSystem.assemble(g, ports, d, solver=model.SolverChoice.ITERATIVE)
    with pytest.raises(SolveError):
        it.grad(FREQS, d_nodes, d_widths)
    it.close()
