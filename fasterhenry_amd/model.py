"""Problem description types: geometry, ports, discretization, solver choice.

A host-side mirror of the reference's library surface (fasterhenry 0.2.0):

* ``Node``/``SegmentDef``/``Geometry`` -- fasterhenry/src/geometry.rs:56-410
* ``Port`` -- fasterhenry/src/mesh.rs:68-94
* ``Subdivision``/``Grading``/``AxisGrading``/``SkinDepthGrading``/
  ``Discretization`` -- fasterhenry/src/solve.rs:191-519
* ``Solver``/``SolverChoice``/``DENSE_PATH_MAX_FILAMENTS`` --
  fasterhenry/src/solve.rs:133-189
* ``skin_depth`` -- fasterhenry/src/solve.rs:556-558

Everything here is plain Python; ``pack`` turns a problem into the plain
arrays that the C ABI (include/fasterhenry_gpu.h) takes.
"""
from __future__ import annotations

import enum
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from .errors import GeometryError, InvalidGrading, MeshError

#: Largest filament count at which SolverChoice.AUTO keeps the dense path
#: (fasterhenry/src/solve.rs:133).
DENSE_PATH_MAX_FILAMENTS = 10_000

#: mu0 = 4 pi 1e-7 H/m (fasterhenry/src/inductance/mod.rs:76).
MU0 = 4.0e-7 * math.pi

# Plain-data layouts shared with include/fasterhenry_gpu.h.
SEGMENT_DTYPE = np.dtype(
    [
        ("a", "<u4"),
        ("b", "<u4"),
        ("width", "<f8"),
        ("height", "<f8"),
        ("sigma", "<f8"),
        ("width_dir", "<f8", (3,)),
        ("has_width_dir", "<u4"),
        ("pad", "<u4"),
    ],
    align=True,
)
GRID_DTYPE = np.dtype(
    [
        ("nw", "<u4"),
        ("nh", "<u4"),
        ("width_ratio", "<f8"),
        ("height_ratio", "<f8"),
        ("graded", "<u4"),
        ("pad", "<u4"),
    ],
    align=True,
)
PORT_DTYPE = np.dtype([("positive", "<u4"), ("negative", "<u4")], align=True)
FILAMENT_DTYPE = np.dtype(
    [
        ("start", "<f8", (3,)),
        ("end", "<f8", (3,)),
        ("length", "<f8"),
        ("dir", "<f8", (3,)),
        ("wdir", "<f8", (3,)),
        ("hdir", "<f8", (3,)),
        ("width", "<f8"),
        ("height", "<f8"),
        ("sigma", "<f8"),
    ],
    align=True,
)
assert SEGMENT_DTYPE.itemsize == 64 and GRID_DTYPE.itemsize == 32
assert FILAMENT_DTYPE.itemsize == 152


@dataclass(frozen=True)
class Node:
    """A point where segments begin, end and connect (geometry.rs:56-83)."""

    x: float
    y: float
    z: float

    def is_finite(self) -> bool:
        return all(math.isfinite(c) for c in (self.x, self.y, self.z))


NodeId = int


@dataclass(frozen=True)
class SegmentDef:
    """A segment between two nodes of a Geometry (geometry.rs:289-330)."""

    a: NodeId
    b: NodeId
    width: float
    height: float
    sigma: float
    width_dir: Optional[Sequence[float]] = None

    def with_width_dir(self, width_dir: Sequence[float]) -> "SegmentDef":
        return SegmentDef(self.a, self.b, self.width, self.height, self.sigma, tuple(width_dir))


class Geometry:
    """Nodes and the segments joining them (geometry.rs:369-410).

    Validation of node references happens as segments are added; the frame
    of each segment (finite end points, positive extents, usable width
    direction) is validated by the native library at assembly, which reports
    GeometryError with the segment index.
    """

    def __init__(self) -> None:
        self.nodes: List[Node] = []
        self.segments: List[SegmentDef] = []

    @classmethod
    def from_parts(cls, nodes: Sequence[Node], segments: Sequence[SegmentDef]) -> "Geometry":
        g = cls()
        for n in nodes:
            g.add_node(n)
        for s in segments:
            g.add_segment(s)
        return g

    def add_node(self, node: Node) -> NodeId:
        if not node.is_finite():
            raise GeometryError(f"node {len(self.nodes)} has a non-finite coordinate",
                                index=len(self.nodes))
        self.nodes.append(node)
        return len(self.nodes) - 1

    def add_segment(self, segment: SegmentDef) -> int:
        for node in (segment.a, segment.b):
            if not 0 <= int(node) < len(self.nodes):
                raise GeometryError(
                    f"segment {len(self.segments)} refers to node {node}, but the geometry has "
                    f"only {len(self.nodes)} nodes",
                    index=len(self.segments),
                )
        self.segments.append(segment)
        return len(self.segments) - 1

    def segment_count(self) -> int:
        return len(self.segments)


@dataclass(frozen=True)
class Port:
    """Terminal pair; current enters at ``positive`` (mesh.rs:68-94)."""

    positive: NodeId
    negative: NodeId
    name: Optional[str] = None


class Solver(enum.Enum):
    """Which solve path runs (solve.rs:143-158)."""

    DENSE = "dense"
    ITERATIVE = "iterative"

    def __str__(self) -> str:
        return self.value


class SolverChoice(enum.Enum):
    """Which path a caller wants before the size is known (solve.rs:169-189)."""

    AUTO = 0
    DENSE = 1
    ITERATIVE = 2

    def resolve(self, filaments: int) -> Solver:
        if self is SolverChoice.DENSE:
            return Solver.DENSE
        if self is SolverChoice.ITERATIVE:
            return Solver.ITERATIVE
        return Solver.DENSE if filaments <= DENSE_PATH_MAX_FILAMENTS else Solver.ITERATIVE


@dataclass(frozen=True)
class Subdivision:
    """Uniform nw x nh cross-section grid (solve.rs:191-209)."""

    nw: int
    nh: int


Subdivision.SINGLE = Subdivision(1, 1)  # type: ignore[attr-defined]


@dataclass(frozen=True)
class Grading:
    """nw x nh grid graded toward the surfaces by ``ratio`` (solve.rs:218-244)."""

    nw: int
    nh: int
    ratio: float

    def validate(self) -> None:
        if not (math.isfinite(self.ratio) and self.ratio >= 1.0):
            raise InvalidGrading(
                f"the grading ratio must be finite and at least 1, got {self.ratio}")


@dataclass(frozen=True)
class AxisGrading:
    """Per-axis graded grid (solve.rs:253-296)."""

    nw: int
    nh: int
    width_ratio: float
    height_ratio: float

    def validate(self) -> None:
        for axis, ratio in (("width", self.width_ratio), ("height", self.height_ratio)):
            if not (math.isfinite(ratio) and ratio >= 1.0):
                raise InvalidGrading(
                    f"the {axis} grading ratio must be finite and at least 1, got {ratio}")


def skin_depth(frequency_hz: float, sigma: float) -> float:
    """delta = 1/sqrt(pi f mu0 sigma) (solve.rs:556-558); inf at f = 0."""
    denom = math.pi * frequency_hz * MU0 * sigma
    return math.inf if denom == 0.0 else 1.0 / math.sqrt(denom)


def _powi(a: float, b: int) -> float:
    """Binary powering, the arithmetic of Rust's f64::powi."""
    recip = b < 0
    b = abs(b)
    r = 1.0
    while True:
        if b & 1:
            r *= a
        b //= 2
        if b == 0:
            break
        a *= a
    return 1.0 / r if recip else r


def graded_weights(count: int, ratio: float) -> List[float]:
    """ratio^min(i, count-1-i) (filament.rs:262-277)."""
    if not (math.isfinite(ratio) and ratio >= 1.0):
        raise InvalidGrading(f"invalid grading ratio {ratio}")
    if count == 0:
        raise InvalidGrading("zero subdivision")
    w = [_powi(ratio, min(i, count - 1 - i)) for i in range(count)]
    if not all(math.isfinite(x) for x in w):
        raise OverflowError("grading overflow")
    return w


def graded_extents(total: float, count: int, ratio: float) -> List[float]:
    """Cell extents of a graded partition (filament.rs:220-225)."""
    w = graded_weights(count, ratio)
    s = 0.0
    for x in w:
        s += x
    return [total * x / s for x in w]


def graded_surface_extent(total: float, count: int, ratio: float) -> float:
    """Extent of the outermost cell (filament.rs:247-251)."""
    w = graded_weights(count, ratio)
    s = 0.0
    for x in w:
        s += x
    return total / s


@dataclass(frozen=True)
class SkinDepthGrading:
    """Counts chosen per segment from the skin depth (solve.rs:324-403)."""

    frequency_hz: float
    target_skin_depths: float
    ratio: float
    max_per_axis: int = 32

    DEFAULT_MAX_PER_AXIS = 32

    def with_max_per_axis(self, max_per_axis: int) -> "SkinDepthGrading":
        return SkinDepthGrading(self.frequency_hz, self.target_skin_depths, self.ratio,
                                max_per_axis)

    def validate(self) -> None:
        if not (math.isfinite(self.ratio) and self.ratio >= 1.0):
            raise InvalidGrading(
                f"the grading ratio must be finite and at least 1, got {self.ratio}")
        if not (math.isfinite(self.target_skin_depths) and self.target_skin_depths > 0.0):
            raise InvalidGrading(
                f"target_skin_depths must be finite and positive, got {self.target_skin_depths}")
        if not (math.isfinite(self.frequency_hz) and self.frequency_hz >= 0.0):
            raise InvalidGrading(
                f"frequency_hz must be finite and non-negative, got {self.frequency_hz}")
        if self.max_per_axis == 0:
            raise InvalidGrading("max_per_axis must be at least 1")

    def count_for(self, extent: float, sigma: float) -> int:
        """solve.rs:388-402"""
        target = self.target_skin_depths * skin_depth(self.frequency_hz, sigma)
        if math.isinf(target) or extent <= target:
            return 1
        for count in range(2, self.max_per_axis + 1):
            try:
                surface = graded_surface_extent(extent, count, self.ratio)
            except OverflowError:
                return count - 1
            if surface <= target:
                return count
        return self.max_per_axis


@dataclass(frozen=True)
class SegmentGrid:
    """The resolved grid of one segment (solve.rs:513-519)."""

    nw: int
    nh: int
    width_ratio: float = 1.0
    height_ratio: float = 1.0
    graded: bool = False


class Discretization:
    """How segments are cut into filaments (solve.rs:409-510).

    Construct with the class methods: ``uniform``, ``per_segment``,
    ``graded``, ``skin_depth``, ``per_segment_graded``.
    """

    def __init__(self, kind: str, value) -> None:
        self.kind = kind
        self.value = value

    @classmethod
    def uniform(cls, nw: int, nh: int) -> "Discretization":
        return cls("uniform", Subdivision(nw, nh))

    @classmethod
    def per_segment(cls, subdivisions: Sequence[Subdivision]) -> "Discretization":
        return cls("per_segment", list(subdivisions))

    @classmethod
    def graded(cls, nw: int, nh: int, ratio: float) -> "Discretization":
        return cls("graded", Grading(nw, nh, ratio))

    @classmethod
    def skin_depth(cls, frequency_hz: float, target_skin_depths: float,
                   ratio: float) -> "Discretization":
        return cls("skin_depth", SkinDepthGrading(frequency_hz, target_skin_depths, ratio))

    @classmethod
    def skin_depth_with(cls, grading: SkinDepthGrading) -> "Discretization":
        return cls("skin_depth", grading)

    @classmethod
    def per_segment_graded(cls, gradings: Sequence[AxisGrading]) -> "Discretization":
        return cls("per_segment_graded", list(gradings))

    def resolve(self, geometry: Geometry) -> List[SegmentGrid]:
        """Per-segment grids (solve.rs:468-510)."""
        n = geometry.segment_count()
        if self.kind == "uniform":
            s = self.value
            return [SegmentGrid(s.nw, s.nh)] * n
        if self.kind == "per_segment":
            if len(self.value) != n:
                raise MeshError(f"got filament counts for {len(self.value)} segments, but the "
                                f"geometry has {n}", kind="SegmentCountMismatch")
            return [SegmentGrid(s.nw, s.nh) for s in self.value]
        if self.kind == "per_segment_graded":
            if len(self.value) != n:
                raise MeshError(f"got filament counts for {len(self.value)} segments, but the "
                                f"geometry has {n}", kind="SegmentCountMismatch")
            out = []
            for g in self.value:
                g.validate()
                out.append(SegmentGrid(g.nw, g.nh, g.width_ratio, g.height_ratio, True))
            return out
        if self.kind == "graded":
            g = self.value
            g.validate()
            return [SegmentGrid(g.nw, g.nh, g.ratio, g.ratio, True)] * n
        if self.kind == "skin_depth":
            g = self.value
            g.validate()
            return [SegmentGrid(g.count_for(s.width, s.sigma), g.count_for(s.height, s.sigma),
                                g.ratio, g.ratio, True) for s in geometry.segments]
        raise ValueError(f"unknown discretization {self.kind}")

    def filament_count(self, geometry: Geometry) -> int:
        """Sum of nw*nh over segments, without cutting (solve.rs:460-466)."""
        return sum(g.nw * g.nh for g in self.resolve(geometry))


@dataclass
class PackedProblem:
    """The plain arrays of include/fasterhenry_gpu.h's fh_problem."""

    nodes: np.ndarray      # (N, 3) float64
    segments: np.ndarray   # SEGMENT_DTYPE
    grids: np.ndarray      # GRID_DTYPE
    ports: np.ndarray      # PORT_DTYPE
    port_names: List[Optional[str]] = field(default_factory=list)

    @property
    def filament_count(self) -> int:
        return int(np.sum(self.grids["nw"].astype(np.int64) * self.grids["nh"]))


def pack(geometry: Geometry, ports: Sequence[Port], discretization: Discretization) -> PackedProblem:
    """Resolve the discretization and lay the problem out as plain arrays."""
    grids_list = discretization.resolve(geometry)
    nodes = np.array([[n.x, n.y, n.z] for n in geometry.nodes], dtype=np.float64).reshape(-1, 3)
    segs = np.zeros(len(geometry.segments), dtype=SEGMENT_DTYPE)
    for i, s in enumerate(geometry.segments):
        segs[i]["a"] = s.a
        segs[i]["b"] = s.b
        segs[i]["width"] = s.width
        segs[i]["height"] = s.height
        segs[i]["sigma"] = s.sigma
        if s.width_dir is not None:
            segs[i]["width_dir"] = s.width_dir
            segs[i]["has_width_dir"] = 1
    grids = np.zeros(len(grids_list), dtype=GRID_DTYPE)
    for i, g in enumerate(grids_list):
        grids[i]["nw"] = g.nw
        grids[i]["nh"] = g.nh
        grids[i]["width_ratio"] = g.width_ratio
        grids[i]["height_ratio"] = g.height_ratio
        grids[i]["graded"] = 1 if g.graded else 0
    prts = np.zeros(len(ports), dtype=PORT_DTYPE)
    for i, p in enumerate(ports):
        for node in (p.positive, p.negative):
            if int(node) < 0:
                raise MeshError(f"port {i} refers to node {node}", kind="UnknownPortNode")
        prts[i]["positive"] = p.positive
        prts[i]["negative"] = p.negative
    return PackedProblem(np.ascontiguousarray(nodes), segs, grids, prts,
                         [p.name for p in ports])
