"""Synthetic geometries the benchmarks and tests run on.

* ``meander`` -- fasterhenry/benches/support.rs:104-127 (the scaling fixture:
  ``meander(49)`` at 2 x 2 is the 9 800-filament case, ``meander(24)`` the
  2 400-filament one).
* ``serpentine`` -- fasterhenry/benches/support.rs:64-83.
* ``busbar_pair`` -- two parallel copper bars shorted at the far end, one port.
* ``coupled_structure`` -- fasterhenry/tests/common/mod.rs:44-100.
* ``spiral_system`` -- fasterhenry/tests/common/mod.rs:102-157, the 2-turn
  square spiral whose published result is L(1 kHz) = 0.637460 nH and
  R(DC) = 2.844828 ohm (docs/validation.md:107-112).
* ``batch_problems`` -- the batched-sweep workload: seeded square spirals and
  busbar pairs of uneven sizes (300 - 4 000 filaments), surface-graded.
* ``AdjointSpiral`` -- the gradient workload: a planar square spiral whose
  turn widths and spacings are the parameters, its discretization frozen at
  the nominal geometry, with the exact node and width velocities.
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np

from .model import (AxisGrading, Discretization, Geometry, Node, Port, SegmentDef,
                    SkinDepthGrading, Subdivision)

COPPER = 5.8e7

#: benches/support.rs:43-44
SUBDIVISION_NW = 2
SUBDIVISION_NH = 2
#: benches/support.rs:89-96
PITCH = 100e-6
WIDTH = 20e-6
THICKNESS = 10e-6


def meander(bars: int) -> Tuple[Geometry, Port]:
    """Square meander of ``bars`` runs, each ``bars`` pitch-long segments."""
    g = Geometry()
    previous = g.add_node(Node(0.0, 0.0, 0.0))
    first = previous

    def add(x: float, y: float) -> None:
        nonlocal previous
        node = g.add_node(Node(x, y, 0.0))
        g.add_segment(SegmentDef(previous, node, WIDTH, THICKNESS, COPPER))
        previous = node

    for bar in range(bars):
        y = bar * PITCH
        for step in range(1, bars + 1):
            along = step if bar % 2 == 0 else bars - step
            add(along * PITCH, y)
        x = (bars if bar % 2 == 0 else 0) * PITCH
        add(x, y + PITCH)
    return g, Port(first, previous)


def meander_filament_count(bars: int) -> int:
    """benches/scaling.rs:99-101"""
    return bars * (bars + 1) * SUBDIVISION_NW * SUBDIVISION_NH


def meander_problem(bars: int):
    g, p = meander(bars)
    return g, [p], Discretization.uniform(SUBDIVISION_NW, SUBDIVISION_NH)


def trace_meander(runs: int = 12, run_length: float = 8e-3, pitch: float = 0.4e-3,
                  width: float = 0.2e-3, thickness: float = 35e-6, segments_per_run: int = 12,
                  nw: int = 5, nh: int = 3, ratio: float = 2.0):
    """A copper trace ``width`` x ``thickness`` folded into ``runs`` parallel
    runs of ``run_length`` at ``pitch``, laid out as meander() lays its
    runs: each run cut into ``segments_per_run`` segments and followed by a
    one-pitch jog (12 x 12 + 12 = 156 segments); every segment ``nw`` x ``nh``
    graded toward the surfaces by ``ratio`` (2 340 filaments at the
    defaults); one port across the ends."""
    g = Geometry()
    previous = g.add_node(Node(0.0, 0.0, 0.0))
    first = previous

    def add(x: float, y: float) -> None:
        nonlocal previous
        node = g.add_node(Node(x, y, 0.0))
        g.add_segment(SegmentDef(previous, node, width, thickness, COPPER))
        previous = node

    step_len = run_length / segments_per_run
    for run in range(runs):
        y = run * pitch
        for step in range(1, segments_per_run + 1):
            along = step if run % 2 == 0 else segments_per_run - step
            add(along * step_len, y)
        x = (segments_per_run if run % 2 == 0 else 0) * step_len
        add(x, y + pitch)
    return g, [Port(first, previous, "trace")], Discretization.graded(nw, nh, ratio)


def serpentine(bars: int) -> Tuple[Geometry, Port]:
    """benches/support.rs:64-83"""
    length, pitch, width, thickness = 2e-3, 60e-6, 20e-6, 10e-6
    g = Geometry()
    previous = g.add_node(Node(0.0, 0.0, 0.0))
    first = previous
    for bar in range(bars):
        y = bar * pitch
        x_far = length if bar % 2 == 0 else 0.0
        for (x, yy) in ((x_far, y), (x_far, y + pitch)):
            node = g.add_node(Node(x, yy, 0.0))
            g.add_segment(SegmentDef(previous, node, width, thickness, COPPER))
            previous = node
    return g, Port(first, previous)


def busbar_pair(segments_per_bar: int = 10, nw: int = 8, nh: int = 4):
    """Two parallel copper bars, 10 mm x 1.0 mm x 0.1 mm, 0.5 mm edge gap,
    shorted at the far end by a single-filament strap; one port across the
    near ends.  640 bar filaments + 1 strap at the defaults."""
    length, width, height, gap = 10e-3, 1.0e-3, 0.1e-3, 0.5e-3
    g = Geometry()
    ya, yb = 0.0, width + gap
    a_nodes = [g.add_node(Node(length * k / segments_per_bar, ya, 0.0))
               for k in range(segments_per_bar + 1)]
    b_nodes = [g.add_node(Node(length * k / segments_per_bar, yb, 0.0))
               for k in range(segments_per_bar + 1)]
    subs: List[Subdivision] = []
    for k in range(segments_per_bar):
        g.add_segment(SegmentDef(a_nodes[k], a_nodes[k + 1], width, height, COPPER))
        subs.append(Subdivision(nw, nh))
    for k in range(segments_per_bar):
        g.add_segment(SegmentDef(b_nodes[k + 1], b_nodes[k], width, height, COPPER))
        subs.append(Subdivision(nw, nh))
    g.add_segment(SegmentDef(a_nodes[-1], b_nodes[-1], width, height, COPPER))
    subs.append(Subdivision.SINGLE)
    return g, [Port(a_nodes[0], b_nodes[0])], Discretization.per_segment(subs)


def coupled_structure():
    """tests/common/mod.rs:44-100: 11 nodes, 8 segments, 41 filaments, a ring."""
    points = [
        [0.0, 0.0, 0.0], [4e-3, 0.0, 0.0], [4e-3, 2e-3, 0.0],
        [0.5e-3, 0.0, 0.3e-3], [3.5e-3, 0.0, 0.3e-3],
        [0.0, -1e-3, 0.0], [3e-3, -1e-3, 0.0], [3e-3, -3e-3, 0.0], [0.0, -3e-3, 0.0],
        [1e-3, 1e-3, 0.0], [3e-3, 1e-3, 0.0],
    ]
    segments = [
        (0, 1, 0.4e-3, 0.1e-3), (1, 2, 0.4e-3, 0.1e-3), (3, 4, 0.2e-3, 0.05e-3),
        (5, 6, 0.2e-3, 0.1e-3), (6, 7, 0.2e-3, 0.1e-3), (7, 8, 0.2e-3, 0.1e-3),
        (8, 5, 0.2e-3, 0.1e-3), (9, 10, 0.2e-3, 0.1e-3),
    ]
    g = Geometry()
    for p in points:
        g.add_node(Node(*p))
    for a, b, w, h in segments:
        g.add_segment(SegmentDef(a, b, w, h, COPPER))
    ports = [Port(0, 2, "A"), Port(4, 3, "B"), Port(5, 7, "ring diagonal")]
    subs = [Subdivision(2, 2)] * len(segments)
    subs[0] = Subdivision(4, 3)
    subs[1] = Subdivision(3, 3)
    subs[2] = Subdivision(3, 1)
    subs[5] = Subdivision.SINGLE
    return g, ports, Discretization.per_segment(subs)


def square_spiral_segments(turns: int, start_len_um: float, pitch_um: float):
    """tests/common/mod.rs:123-142"""
    directions = [(1.0, 0.0), (0.0, 1.0), (-1.0, 0.0), (0.0, -1.0)]
    position = (0.0, 0.0)
    length = start_len_um
    out = []
    for k in range(turns * 4):
        d = directions[k % 4]
        end = (position[0] + d[0] * length, position[1] + d[1] * length)
        out.append((position, end))
        position = end
        if k % 2 == 1:
            length += pitch_um
    return out


def spiral_system():
    """tests/common/mod.rs:144-157: 2 turns, 60 um start, 15 um pitch, 2x2 um
    copper, 4 x 4 filaments per segment (128 filaments)."""
    um, w_um, t_um = 1e-6, 2.0, 2.0
    segs = square_spiral_segments(2, 60.0, 15.0)
    nodes = [Node(0.0, 0.0, t_um / 2.0 * um)]
    for (_, end) in segs:
        nodes.append(Node(end[0] * um, end[1] * um, t_um / 2.0 * um))
    defs = [SegmentDef(i, i + 1, w_um * um, t_um * um, COPPER) for i in range(len(segs))]
    g = Geometry.from_parts(nodes, defs)
    return g, [Port(0, len(segs), "spiral")], Discretization.per_segment(
        [Subdivision(4, 4)] * len(segs))


def log_frequencies(f0: float, f1: float, count: int) -> np.ndarray:
    """f_k = f0 (f1/f0)^(k/(count-1))."""
    if count == 1:
        return np.array([f0])
    k = np.arange(count, dtype=np.float64)
    return f0 * (f1 / f0) ** (k / (count - 1))


def lcg(seed: int):
    """Uniform deviates in [0, 1) (fasterhenry/src/pfft/cloud_draws.rs:36-44)."""
    state = seed & 0xFFFFFFFFFFFFFFFF

    def draw() -> float:
        nonlocal state
        state = (state * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        return (state >> 11) / float(1 << 53)

    return draw


def cloud_side(n: int) -> float:
    """cloud_draws.rs:54-56"""
    return 10e-3 * float(np.cbrt(n / 400.0))


def random_bars(n: int, side: float, manhattan: bool, seed: int):
    """cloud_draws.rs:69-95: bars of length 0.5-1.5 mm in a cube; Rust
    evaluates an assignment's value before its place, so the sign draw
    precedes the axis draw."""
    r = lcg(seed)
    out = []
    for _ in range(n):
        a = [side * r(), side * r(), side * r()]
        length = 1e-3 * (0.5 + r())
        if manhattan:
            d = [0.0, 0.0, 0.0]
            sign = -1.0 if r() < 0.5 else 1.0
            d[min(int(r() * 3.0), 2)] = sign
        else:
            v = [r() - 0.5, r() - 0.5, r() - 0.5]
            nrm = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
            d = [c / nrm for c in v]
        b = [a[k] + length * d[k] for k in range(3)]
        width = 0.1e-3 + 0.2e-3 * r()
        height = 0.05e-3 + 0.1e-3 * r()
        out.append((a, b, width, height))
    return out


def random_cloud_problem(n: int, manhattan: bool, seed: int, side: float = None):
    """The bars of random_bars as a geometry of single-filament segments
    (benches/support.rs:147-163 random_set), with one port on bar 0."""
    side = cloud_side(n) if side is None else side
    g = Geometry()
    for (a, b, w, h) in random_bars(n, side, manhattan, seed):
        na = g.add_node(Node(*a))
        nb = g.add_node(Node(*b))
        g.add_segment(SegmentDef(na, nb, w, h, COPPER))
    return g, [Port(0, 1)], Discretization.uniform(1, 1)


def graded_spiral(turns: int, width: float, spacing: float, thickness: float, nw: int, nh: int,
                  split: int = 1, ratio: float = 2.0):
    """A square spiral of ``turns`` turns in the plane z = thickness / 2,
    innermost side 4 (width + spacing) long, pitch width + spacing, each side
    cut into ``split`` equal segments; one port across its ends; every
    segment cut nw x nh, graded toward the surfaces by ``ratio``."""
    pitch = width + spacing
    sides = square_spiral_segments(turns, 4.0 * pitch, pitch)
    nodes = [Node(0.0, 0.0, 0.5 * thickness)]
    for (start, end) in sides:
        for q in range(1, split + 1):
            t = q / split
            nodes.append(Node(start[0] + t * (end[0] - start[0]), start[1] + t * (end[1] - start[1]),
                              0.5 * thickness))
    defs = [SegmentDef(i, i + 1, width, thickness, COPPER) for i in range(len(nodes) - 1)]
    g = Geometry.from_parts(nodes, defs)
    return g, [Port(0, len(nodes) - 1, "spiral")], Discretization.graded(nw, nh, ratio)


def graded_busbar_pair(width: float, spacing: float, thickness: float, nw: int, nh: int,
                       segments_per_bar: int = 8, length: float = 5e-3, ratio: float = 2.0):
    """Two parallel bars ``length`` long, ``spacing`` apart edge to edge,
    shorted at the far end by a strap of the same section; one port across the
    near ends; every segment cut nw x nh, graded toward the surfaces."""
    g = Geometry()
    ya, yb = 0.0, width + spacing
    z = 0.5 * thickness
    a = [g.add_node(Node(length * k / segments_per_bar, ya, z)) for k in range(segments_per_bar + 1)]
    b = [g.add_node(Node(length * k / segments_per_bar, yb, z)) for k in range(segments_per_bar + 1)]
    for k in range(segments_per_bar):
        g.add_segment(SegmentDef(a[k], a[k + 1], width, thickness, COPPER))
    for k in range(segments_per_bar):
        g.add_segment(SegmentDef(b[k + 1], b[k], width, thickness, COPPER))
    g.add_segment(SegmentDef(a[-1], b[-1], width, thickness, COPPER))
    return g, [Port(a[0], b[0])], Discretization.graded(nw, nh, ratio)


def _cross_section(width: float, thickness: float):
    """Surface-graded cells about as wide as tall: 2 or 3 across the
    thickness, 2 - 10 across the width."""
    nh = 2 if thickness < 40e-6 else 3
    nw = max(2, min(10, int(round(nh * width / thickness))))
    return nw, nh


def batch_problems(count: int = 48, seed: int = 48):
    """The batched-sweep workload: ``count`` geometries from the LCG of
    fasterhenry/src/pfft/cloud_draws.rs:36-44.  Five draws per geometry, in
    this order: a turn draw r, width U(50, 500) um, spacing U(50, 300) um,
    thickness U(17, 70) um, a size draw u.  Even index: a square spiral of
    2 + floor(7 r) turns; odd index: a busbar pair.  Each segment is cut
    nw x nh (graded, ratio 2; see _cross_section), and the sides are split
    into as many segments as bring the filament count closest to
    300 + 3 700 u."""
    r = lcg(seed)
    out = []
    for k in range(count):
        r_turn = r()
        width = 50e-6 + 450e-6 * r()
        spacing = 50e-6 + 250e-6 * r()
        thickness = 17e-6 + 53e-6 * r()
        target = 300 + 3700 * r()
        nw, nh = _cross_section(width, thickness)
        if k % 2 == 0:
            turns = 2 + min(int(7 * r_turn), 6)
            split = max(1, int(round(target / (4 * turns * nw * nh))))
            out.append(graded_spiral(turns, width, spacing, thickness, nw, nh, split))
        else:
            per_bar = max(1, int(round((target / (nw * nh) - 1) / 2)))
            out.append(graded_busbar_pair(width, spacing, thickness, nw, nh, per_bar))
    return out


class AdjointSpiral:
    """A planar square spiral parameterized by its turn widths w_k and
    spacings s_k (p = [w_0 .. w_{T-1}, s_0 .. s_{T-1}]), outermost turn first.

    The outer edge is a square of side ``outer`` centred on the origin, in
    the plane z = thickness / 2.  Turn k's centreline sits a_k = sum_{i<k}
    (w_i + s_i) + w_k / 2 in from that square; its top, right and bottom sides
    run at that offset and its left side climbs to the next turn's top side
    (the last one to a_{T-1} + w_{T-1} + s_{T-1}, so every spacing counts).
    One port across the outer and inner ends.

    The discretization is frozen at the nominal parameters: side j is cut
    into ``splits[j]`` segments (about equal lengths, giving about
    ``filaments`` filaments) and turn k's segments into the skin-depth counts
    of SkinDepthGrading(frequency, target, ratio) for w_k x thickness, as
    per-segment graded grids.  Node positions are affine in p, so
    ``directions()`` returns exact velocities."""

    def __init__(self, turns: int = 6, outer: float = 2e-3, width: float = 100e-6,
                 spacing: float = 50e-6, thickness: float = 35e-6, frequency: float = 1e9,
                 target: float = 0.5, ratio: float = 2.0, filaments: int = 6000):
        self.turns = turns
        self.outer = outer
        self.thickness = thickness
        self.nominal = np.array([width] * turns + [spacing] * turns, dtype=np.float64)
        grading = SkinDepthGrading(frequency, target, ratio)
        self.counts = [(grading.count_for(width, COPPER), grading.count_for(thickness, COPPER))] * turns
        self.ratio = ratio
        corners = self.corners(self.nominal)
        lengths = [math.hypot(b[0] - a[0], b[1] - a[1]) for a, b in zip(corners[:-1], corners[1:])]
        per = [self.counts[j // 4][0] * self.counts[j // 4][1] for j in range(len(lengths))]
        segments = filaments / (sum(per) / len(per))
        step = sum(lengths) / segments
        self.splits = [max(1, int(round(L / step))) for L in lengths]

    @property
    def n_params(self) -> int:
        return 2 * self.turns

    def corners(self, params) -> List[Tuple[float, float]]:
        w = [float(x) for x in params[:self.turns]]
        sp = [float(x) for x in params[self.turns:]]
        h = 0.5 * self.outer
        a, acc = [], 0.0
        for k in range(self.turns):
            a.append(acc + 0.5 * w[k])
            acc += w[k] + sp[k]
        a.append(a[-1] + w[-1] + sp[-1])
        pts = [(-h + a[0], h - a[0])]
        for k in range(self.turns):
            pts += [(h - a[k], h - a[k]), (h - a[k], -h + a[k]), (-h + a[k], -h + a[k]),
                    (-h + a[k], h - a[k + 1])]
        return pts

    def node_positions(self, params) -> np.ndarray:
        pts = self.corners(params)
        out = [(pts[0][0], pts[0][1])]
        for j, (a, b) in enumerate(zip(pts[:-1], pts[1:])):
            n = self.splits[j]
            for q in range(1, n + 1):
                t = q / n
                out.append((a[0] + t * (b[0] - a[0]), a[1] + t * (b[1] - a[1])))
        xyz = np.zeros((len(out), 3))
        xyz[:, :2] = out
        xyz[:, 2] = 0.5 * self.thickness
        return xyz

    def segment_turns(self) -> List[int]:
        return [j // 4 for j, n in enumerate(self.splits) for _ in range(n)]

    def problem(self, params=None):
        """(geometry, ports, discretization) at ``params`` (default nominal)."""
        p = self.nominal if params is None else np.asarray(params, dtype=np.float64)
        xyz = self.node_positions(p)
        nodes = [Node(*map(float, r)) for r in xyz]
        turn = self.segment_turns()
        defs = [SegmentDef(i, i + 1, float(p[turn[i]]), self.thickness, COPPER)
                for i in range(len(turn))]
        g = Geometry.from_parts(nodes, defs)
        grids = [AxisGrading(self.counts[k][0], self.counts[k][1], self.ratio, self.ratio)
                 for k in turn]
        return g, [Port(0, len(nodes) - 1, "spiral")], Discretization.per_segment_graded(grids)

    def directions(self) -> Tuple[np.ndarray, np.ndarray]:
        """(d_nodes (P, n_nodes, 3), d_widths (P, n_segments)): node positions
        are affine in p, so a unit step gives the exact velocity."""
        base = self.node_positions(self.nominal)
        P = self.n_params
        d_nodes = np.zeros((P,) + base.shape)
        for q in range(P):
            e = self.nominal.copy()
            e[q] += 1.0
            d_nodes[q] = self.node_positions(e) - base
        turn = np.array(self.segment_turns())
        d_widths = np.zeros((P, len(turn)))
        for k in range(self.turns):
            d_widths[k, turn == k] = 1.0
        return d_nodes, d_widths
