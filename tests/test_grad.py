"""Host side of the gradients: the filament tangents of fh_discretize_tangent
against central differences of the oracle's discretization, and the
gradient workload's geometry (exact affine node velocities, frozen counts)."""
import numpy as np
import pytest

import oracle
from fasterhenry_amd import api, fixtures, model

FIELDS = ["start", "end", "length", "dir", "wdir", "hdir", "width", "height"]


def _fd_tangent(spiral, q, rel=1e-6):
    h = rel * spiral.nominal[q]
    plus, minus = spiral.nominal.copy(), spiral.nominal.copy()
    plus[q] += h
    minus[q] -= h
    fp = oracle.discretize(model.pack(*spiral.problem(plus)))
    fm = oracle.discretize(model.pack(*spiral.problem(minus)))
    return {k: (np.asarray(fp[k]) - np.asarray(fm[k])) / (2 * h) for k in FIELDS}


@pytest.mark.parametrize("turns", [2, 3])
def test_tangent_matches_central_differences(turns):
    s = fixtures.AdjointSpiral(turns=turns, filaments=300 * turns)
    packed = model.pack(*s.problem())
    d_nodes, d_widths = s.directions()
    for q in range(s.n_params):
        t = api.discretize_tangent(packed, d_nodes[q], d_widths[q])
        fd = _fd_tangent(s, q)
        for k in FIELDS:
            a = np.asarray(t[k])
            # positions and sizes: relative to the largest derivative; unit
            # vectors: absolute (their FD carries the rounding of the node
            # coordinates divided by the step)
            scale = 1.0 if k in ("dir", "wdir", "hdir") else max(np.abs(fd[k]).max(), 1e-30)
            assert np.abs(a - fd[k]).max() <= 1e-6 * scale, (q, k)
        assert np.all(np.asarray(t["sigma"]) == 0.0)


def test_tangent_of_a_rotation_and_of_a_cross_section_change():
    # a single skew bar: rotating its end node about the start turns dir and
    # the frame; changing width and height scales the cells
    g = model.Geometry.from_parts([model.Node(0.0, 0.0, 0.0), model.Node(1e-3, 2e-4, 3e-4)],
                                  [model.SegmentDef(0, 1, 1e-4, 5e-5, fixtures.COPPER)])
    d = model.Discretization.graded(3, 2, 2.0)
    packed = model.pack(g, [], d)
    dn = np.array([[0.0, 0.0, 0.0], [-2e-4, 1e-3, 0.5e-4]])
    dw, dh = np.array([0.3]), np.array([-0.2])
    t = api.discretize_tangent(packed, dn, dw, dh)
    h = 1e-7
    f = []
    for sgn in (1.0, -1.0):
        nodes = packed.nodes + sgn * h * dn
        segs = packed.segments.copy()
        segs["width"] += sgn * h * dw
        segs["height"] += sgn * h * dh
        f.append(oracle.discretize(model.PackedProblem(nodes, segs, packed.grids, packed.ports)))
    for k in FIELDS:
        fd = (np.asarray(f[0][k]) - np.asarray(f[1][k])) / (2 * h)
        scale = max(np.abs(fd).max(), 1e-30)
        assert np.abs(np.asarray(t[k]) - fd).max() <= 1e-6 * scale, k


def test_adjoint_spiral_layout():
    s = fixtures.AdjointSpiral()
    g, ports, d = s.problem()
    packed = model.pack(g, ports, d)
    # skin-depth counts at 1 GHz (0.5 delta, ratio 2) for 100 x 35 um, frozen
    assert s.counts[0] == (12, 9)
    assert 5500 <= packed.filament_count <= 6500
    assert s.n_params == 12
    # edge gaps between neighbouring turns are the spacings
    xyz = s.node_positions(s.nominal)
    assert np.isclose(xyz[0, 0], -1e-3 + 50e-6) and np.isclose(xyz[0, 1], 1e-3 - 50e-6)
    corners = s.corners(s.nominal)
    top0, top1 = corners[1][1], corners[5][1]
    assert np.isclose(top0 - top1, 0.5 * 100e-6 + 50e-6 + 0.5 * 100e-6)
    # velocities are exact for the affine node map
    d_nodes, d_widths = s.directions()
    p = s.nominal.copy()
    p[3] += 1e-6
    p[8] -= 2e-6
    moved = s.node_positions(p)
    assert np.allclose(moved, xyz + 1e-6 * d_nodes[3] - 2e-6 * d_nodes[8], rtol=0, atol=1e-18)
    assert np.array_equal(d_widths.sum(axis=0), np.ones(len(packed.segments)))
    assert np.all(d_widths[6:] == 0.0)
