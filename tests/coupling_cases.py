"""Fixtures of fasterhenry/tests/coupling.rs and coupling.rs's unit tests:
two nearly closed square islands a controlled distance apart, and two
parallel unit traces."""
import math

from fasterhenry_amd import model

COPPER = 5.8e7
SIDE = 10e-3
WIDTH = 0.4e-3
THICKNESS = 35e-6
EXTENT = math.sqrt(2.0) * SIDE
FREQUENCY = 1e9


def island(g, x0):
    gap = SIDE / 20.0
    pts = [(0.0, 0.0), (SIDE, 0.0), (SIDE, SIDE), (0.0, SIDE), (0.0, gap)]
    n = [g.add_node(model.Node(x0 + x, y, 0.0)) for (x, y) in pts]
    for a, b in zip(n[:-1], n[1:]):
        g.add_segment(model.SegmentDef(a, b, WIDTH, THICKNESS, COPPER))
    return model.Port(n[0], n[-1])


def two_islands(separation):
    g = model.Geometry()
    first = island(g, 0.0)
    second = island(g, SIDE + separation)
    return g, [first, second], ["a"] * 4 + ["b"] * 4


def discretization():
    return model.Discretization.uniform(2, 1)


def two_traces(separation):
    g = model.Geometry()
    for y in (0.0, separation):
        a = g.add_node(model.Node(0.0, y, 0.0))
        b = g.add_node(model.Node(1.0, y, 0.0))
        g.add_segment(model.SegmentDef(a, b, 0.1, 0.1, COPPER))
    return g
