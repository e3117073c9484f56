"""Coupling truncation: segment groups and the group pairs whose mutual
inductance is kept (fasterhenry/src/coupling.rs).

The mirror of the reference's ``Coupling`` (names, validation, the
transitivity check, truncation warnings) lives on the host side; the device
receives the resolved mask (``fh_coupling`` / ``fh_pair_mask``) and never
evaluates a truncated pair (inductance/batch.rs:305-353).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .errors import SolveError

#: coupling.rs:90
DEFAULT_GROUP = ""


class CouplingError(SolveError):
    """Why a coupling does not fit its geometry (coupling.rs:93-128)."""


class UnknownGroup(CouplingError):
    pass


class SegmentCountMismatch(CouplingError):
    pass


class NonTransitiveCoupling(CouplingError):
    pass


def _rust_exp3(x: float) -> str:
    """Rust's ``{:.3e}``: three decimals, exponent without '+' or padding."""
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    mant, exp = f"{x:.3e}".split("e")
    return f"{mant}e{int(exp)}"


@dataclass
class TruncationWarning:
    """A truncated group pair closer than one extent apart (coupling.rs:130-168)."""

    groups: Tuple[str, str]
    separation: float
    extent: float

    def __str__(self) -> str:
        def name(g: str) -> str:
            return "(default)" if g == "" else f"'{g}'"
        ratio = self.separation / self.extent if self.extent > 0.0 else math.inf
        ratio_s = "inf" if math.isinf(ratio) else f"{ratio:.2f}"
        return (f"mutual inductance between groups {name(self.groups[0])} and {name(self.groups[1])} "
                f"is truncated, but they are only {_rust_exp3(self.separation)} m apart — {ratio_s}x "
                f"their {_rust_exp3(self.extent)} m extent; truncation is justified from about 10x")


@dataclass
class PairMask:
    """PairMask (inductance/batch.rs:37-110)."""

    group: np.ndarray   # int32 per filament
    groups: int
    coupled: np.ndarray  # uint8, groups x groups

    @classmethod
    def from_groups(cls, group: Sequence[int], groups: int, coupled) -> "PairMask":
        table = np.zeros((groups, groups), dtype=np.uint8)
        for a in range(groups):
            for b in range(a, groups):
                keep = a == b or coupled(a, b) or coupled(b, a)
                table[a, b] = table[b, a] = keep
        return cls(np.asarray(group, dtype=np.int32), groups, table)

    def keeps(self, i: int, j: int) -> bool:
        if i >= len(self.group) or j >= len(self.group):
            return True
        flat = int(self.group[i]) * self.groups + int(self.group[j])
        return bool(self.coupled.reshape(-1)[flat]) if flat < self.groups * self.groups else True

    def total_pairs(self) -> int:
        n = len(self.group)
        return n * (n + 1) // 2

    def kept_pairs(self) -> int:
        sizes = [0] * self.groups
        for g in self.group:
            if 0 <= g < self.groups:
                sizes[g] += 1
        dropped = sum(sizes[a] * sizes[b] for a in range(self.groups) for b in range(a + 1, self.groups)
                      if not self.coupled[a, b])
        return self.total_pairs() - dropped


class _Box:
    def __init__(self):
        self.lo = [math.inf] * 3
        self.hi = [-math.inf] * 3

    def add(self, p):
        for k in range(3):
            self.lo[k] = min(self.lo[k], p[k])
            self.hi[k] = max(self.hi[k], p[k])

    def empty(self) -> bool:
        return any(self.lo[k] > self.hi[k] for k in range(3))

    def diagonal(self) -> float:
        if self.empty():
            return 0.0
        return math.sqrt(sum((self.hi[k] - self.lo[k]) ** 2 for k in range(3)))

    def gap(self, other: "_Box") -> float:
        if self.empty() or other.empty():
            return math.inf
        s = 0.0
        for k in range(3):
            apart = max(self.lo[k] - other.hi[k], other.lo[k] - self.hi[k])
            s += max(apart, 0.0) ** 2
        return math.sqrt(s)


class Coupling:
    """Segment group names and the coupled group pairs (coupling.rs:170-353).
    ``Coupling.all_pairs()`` couples everything (the default);
    ``Coupling.truncated(groups)`` keeps each group with itself only, until
    pairs are declared with ``coupled``."""

    def __init__(self, segment_groups: Optional[List[str]] = None,
                 couples: Optional[List[Tuple[str, str]]] = None):
        self.segment_groups = list(segment_groups or [])
        self._couples = couples

    @classmethod
    def all_pairs(cls) -> "Coupling":
        return cls()

    @classmethod
    def truncated(cls, segment_groups: Sequence[str]) -> "Coupling":
        return cls(list(segment_groups), [])

    def coupled(self, a: str, b: str) -> "Coupling":
        couples = list(self._couples or [])
        couples.append((a, b))
        return Coupling(self.segment_groups, couples)

    def is_all_pairs(self) -> bool:
        return self._couples is None

    def is_trivial(self) -> bool:
        return not self.segment_groups and self._couples is None

    def group_of(self, index: int) -> str:
        return self.segment_groups[index] if index < len(self.segment_groups) else DEFAULT_GROUP

    def couples(self, a: str, b: str) -> bool:
        if self._couples is None or a == b:
            return True
        return any((x == a and y == b) or (x == b and y == a) for (x, y) in self._couples)

    def validate(self, segment_count: int) -> None:
        self._resolve(segment_count)

    def _resolve(self, segment_count: int):
        """(group per segment, names, coupled table) -- coupling.rs:302-353."""
        if len(self.segment_groups) > segment_count:
            raise SegmentCountMismatch(
                f"coupling names {len(self.segment_groups)} segment groups, but the geometry has "
                f"{segment_count} segments", expected=segment_count, got=len(self.segment_groups))
        index: Dict[str, int] = {}
        names: List[str] = []
        of_segment = []
        for q in range(segment_count):
            name = self.group_of(q)
            if name not in index:
                index[name] = len(names)
                names.append(name)
            of_segment.append(index[name])
        g = len(names)
        table = np.eye(g, dtype=bool)
        if self._couples is not None:
            for (a, b) in self._couples:
                for nm in (a, b):
                    if nm not in index:
                        raise UnknownGroup(f"coupling declares group '{nm}', which no segment belongs to",
                                           name=nm)
                table[index[a], index[b]] = table[index[b], index[a]] = True
        else:
            table[:] = True
        _check_transitive(table, names)
        return of_segment, names, table

    def pair_mask(self, filaments_per_segment: Sequence[int]) -> PairMask:
        of_segment, names, table = self._resolve(len(filaments_per_segment))
        group = np.repeat(np.asarray(of_segment, dtype=np.int32),
                          np.asarray(filaments_per_segment, dtype=np.int64))
        return PairMask.from_groups(group, len(names), lambda a, b: bool(table[a, b]))

    def resolved(self, segment_count: int) -> Tuple[np.ndarray, int, np.ndarray]:
        """(int32 group per segment, groups, uint8 symmetric table)."""
        of_segment, names, table = self._resolve(segment_count)
        return (np.asarray(of_segment, dtype=np.int32), len(names),
                np.ascontiguousarray(table.astype(np.uint8)))

    def truncation_warnings(self, geometry) -> List[TruncationWarning]:
        """coupling.rs:274-299: every truncated group pair whose bounding
        boxes lie closer than the larger of their diagonals."""
        of_segment, names, table = self._resolve(geometry.segment_count())
        if self.is_all_pairs():
            return []
        boxes = [_Box() for _ in names]
        for q, seg in enumerate(geometry.segments):
            a, b = geometry.nodes[seg.a], geometry.nodes[seg.b]
            boxes[of_segment[q]].add((a.x, a.y, a.z))
            boxes[of_segment[q]].add((b.x, b.y, b.z))
        out = []
        for a in range(len(names)):
            for b in range(a + 1, len(names)):
                if table[a, b]:
                    continue
                extent = max(boxes[a].diagonal(), boxes[b].diagonal())
                separation = boxes[a].gap(boxes[b])
                if separation < extent:
                    out.append(TruncationWarning((names[a], names[b]), separation, extent))
        return out


def _check_transitive(table: np.ndarray, names: List[str]) -> None:
    """coupling.rs:361-421: each connected component of the declared pairs
    must be a clique."""
    g = len(names)
    if g < 3:
        return
    parent = list(range(g))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for a in range(g):
        for b in range(a + 1, g):
            if table[a, b]:
                ra, rb = find(a), find(b)
                if ra != rb:
                    parent[ra] = rb
    for a in range(g):
        for b in range(a + 1, g):
            if not table[a, b] and find(a) == find(b):
                raise NonTransitiveCoupling(
                    f"coupling is not transitive: groups '{names[a]}' and '{names[b]}' are only "
                    f"reachable through a third group, but are not declared coupled directly — "
                    f"declare the whole clique", groups=(names[a], names[b]))
