"""Geometries and known answers for the filament-current tests
(fasterhenry/tests/currents.rs): a straight bar, a closed square ring with
the port across its first side, two separate one-filament bars, and the
textbook skin-effect slab profile."""
import cmath
import math

from fasterhenry_amd import model

COPPER = 5.8e7


def bar(length, width, height):
    g = model.Geometry()
    a = g.add_node(model.Node(0.0, 0.0, 0.0))
    b = g.add_node(model.Node(length, 0.0, 0.0))
    g.add_segment(model.SegmentDef(a, b, width, height, COPPER))
    return g, [model.Port(a, b)]


def ring():
    g = model.Geometry()
    s = 5e-3
    n = [g.add_node(model.Node(x, y, 0.0)) for (x, y) in [(0.0, 0.0), (s, 0.0), (s, s), (0.0, s)]]
    for i in range(4):
        g.add_segment(model.SegmentDef(n[i], n[(i + 1) % 4], 1e-3, 0.5e-3, COPPER))
    return g, [model.Port(n[0], n[1])]


def two_bars():
    g = model.Geometry()
    n = [g.add_node(model.Node(i * 1e-3, 0.0, 0.0)) for i in range(4)]
    g.add_segment(model.SegmentDef(n[0], n[1], 1e-4, 1e-4, COPPER))
    g.add_segment(model.SegmentDef(n[2], n[3], 1e-4, 1e-4, COPPER))
    return g, [model.Port(n[0], n[1]), model.Port(n[2], n[3])]


def slab_profile(width, t, delta):
    """|J(z)| of a wide slab carrying 1 A, z from the mid-plane, and the
    surface value: J(z) = k cosh(kz) / (2 w sinh(k t / 2)), k = (1 + j) / delta."""
    k = complex(1.0, 1.0) / delta

    def profile(z):
        return abs(k / (2.0 * width) * cmath.cosh(k * z) / cmath.sinh(k * (0.5 * t)))

    surface = abs(k / (2.0 * width) / cmath.tanh(k * (0.5 * t)))
    return profile, surface


def slab_frequency(delta):
    return 1.0 / (math.pi * model.MU0 * COPPER * delta * delta)
