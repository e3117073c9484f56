"""FastHenry `.inp` decks (fasterhenry-cli/src/inp.rs, problem.rs, lib.rs).

``parse`` reads a deck through the native reader (fh_deck_parse, the
conductor subset) into the reference's ``Deck``: title, geometry, ports,
discretization, coupling, frequencies.  ``run`` is the CLI's solve of a deck
(fasterhenry-cli/src/lib.rs:89-107, 183-202): the solver choice (a `.couples`
truncation stays dense), assembly with the coupling, the sweep.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import native
from .coupling import Coupling
from .errors import InvalidArgument
from .model import (AxisGrading, Discretization, Geometry, Node, Port, SegmentDef, SolverChoice,
                    Subdivision)


@dataclass
class Deck:
    """A parsed deck (inp.rs:600-617)."""

    title: Optional[str]
    geometry: Geometry
    ports: List[Port]
    discretization: Discretization
    coupling: Coupling = field(default_factory=Coupling.all_pairs)
    frequencies: List[float] = field(default_factory=list)


def parse(text: str) -> Deck:
    """Parse a deck (inp.rs:3447-3481); a rejected line raises DeckError
    with ``index0`` = the line."""
    L = native.lib()
    raw = text.encode("utf-8")
    h = C.c_void_p()
    native.check(L.fh_deck_parse(raw, len(raw), C.byref(h)))
    try:
        prob = native.Problem()
        native.check(L.fh_deck_problem(h, C.byref(prob)))
        nf = C.c_size_t()
        fp = C.POINTER(C.c_double)()
        disc = C.c_int32()
        ng = C.c_size_t()
        native.check(L.fh_deck_info(h, C.byref(nf), C.byref(fp), C.byref(disc), C.byref(ng)))
        freqs = [float(fp[k]) for k in range(nf.value)]
        nodes = np.ctypeslib.as_array(C.cast(prob.nodes, C.POINTER(C.c_double)), shape=(prob.n_nodes * 3,))
        from .model import GRID_DTYPE, PORT_DTYPE, SEGMENT_DTYPE

        def view(ptr, dtype, count):
            if count == 0:
                return np.zeros(0, dtype=dtype)
            buf = (C.c_char * (count * dtype.itemsize)).from_address(ptr)
            return np.frombuffer(bytes(buf), dtype=dtype).copy()

        segs = view(prob.segments, SEGMENT_DTYPE, prob.n_segments)
        grids = view(prob.grids, GRID_DTYPE, prob.n_segments)
        ports = view(prob.ports, PORT_DTYPE, prob.n_ports)
        g = Geometry()
        for k in range(prob.n_nodes):
            g.add_node(Node(float(nodes[3 * k]), float(nodes[3 * k + 1]), float(nodes[3 * k + 2])))
        for s in segs:
            d = SegmentDef(int(s["a"]), int(s["b"]), float(s["width"]), float(s["height"]), float(s["sigma"]))
            if s["has_width_dir"]:
                d = d.with_width_dir([float(x) for x in s["width_dir"]])
            g.add_segment(d)
        port_list = [Port(int(p["positive"]), int(p["negative"]), L.fh_deck_port_name(h, k).decode())
                     for k, p in enumerate(ports)]
        if disc.value == 2:
            dz = Discretization.per_segment_graded(
                [AxisGrading(int(x["nw"]), int(x["nh"]), float(x["width_ratio"]), float(x["height_ratio"]))
                 for x in grids])
        elif disc.value == 1:
            dz = Discretization.per_segment([Subdivision(int(x["nw"]), int(x["nh"])) for x in grids])
        else:
            dz = Discretization.uniform(int(grids[0]["nw"]), int(grids[0]["nh"]))
        cm = native.CouplingMask()
        native.check(L.fh_deck_coupling(h, C.byref(cm)))
        if cm.coupled:
            names = [L.fh_deck_group_name(h, k).decode() for k in range(cm.groups)]
            group = np.ctypeslib.as_array(C.cast(cm.segment_group, C.POINTER(C.c_int32)),
                                          shape=(cm.n_segments,))
            table = np.ctypeslib.as_array(C.cast(cm.coupled, C.POINTER(C.c_uint8)),
                                          shape=(cm.groups, cm.groups))
            coupling = Coupling.truncated([names[int(x)] for x in group])
            for a in range(cm.groups):
                for b in range(a + 1, cm.groups):
                    if table[a, b]:
                        coupling = coupling.coupled(names[a], names[b])
        else:
            coupling = Coupling.all_pairs()
        title = L.fh_deck_title(h)
        return Deck(title.decode() if title is not None else None, g, port_list, dz, coupling, freqs)
    finally:
        L.fh_deck_destroy(h)


def read(path: str) -> Deck:
    with open(path, "r", encoding="utf-8") as fh:
        return parse(fh.read())


def run(deck: Deck, frequencies=None, solver: SolverChoice = SolverChoice.AUTO,
        precision: str = "fp32_refined"):
    """Solve a deck as the CLI does (fasterhenry-cli/src/lib.rs:89-107,
    183-202): a `.couples` truncation stays dense (and refuses the iterative
    solver), then assemble with the coupling and sweep ``frequencies`` (the
    deck's sweep by default).  Returns the SweepResult."""
    from .api import MeshSystem

    freqs = deck.frequencies if frequencies is None else list(frequencies)
    if not deck.ports:
        raise InvalidArgument("the deck has no ports")
    if not freqs:
        raise InvalidArgument("the deck has no frequencies")
    if deck.coupling.is_all_pairs():
        system = MeshSystem.assemble(deck.geometry, deck.ports, deck.discretization, precision, solver)
    else:
        system = MeshSystem.assemble_with_coupling(deck.geometry, deck.ports, deck.discretization,
                                                   deck.coupling, precision, solver)
    try:
        return system.sweep(freqs)
    finally:
        system.close()


# ---------------------------------------------------------------------------
# The CLI's JSON problem (fasterhenry-cli/src/problem.rs:13-36): serde's
# layout of Geometry, Port, Discretization (snake_case, externally tagged),
# Coupling and the sweep; unknown fields are rejected.

def _check_keys(obj: dict, allowed, what: str) -> None:
    for key in obj:
        if key not in allowed:
            raise InvalidArgument(f"unknown field `{key}` in {what}")


def problem_from_json(text: str) -> Deck:
    import json

    from .model import Grading, SkinDepthGrading

    doc = json.loads(text)
    _check_keys(doc, ("geometry", "ports", "discretization", "coupling", "frequencies_hz"), "problem")
    geo = doc["geometry"]
    _check_keys(geo, ("nodes", "segments"), "geometry")
    g = Geometry()
    for n in geo["nodes"]:
        _check_keys(n, ("x", "y", "z"), "node")
        g.add_node(Node(float(n["x"]), float(n["y"]), float(n["z"])))
    for s in geo["segments"]:
        _check_keys(s, ("a", "b", "width", "height", "sigma", "width_dir"), "segment")
        d = SegmentDef(int(s["a"]), int(s["b"]), float(s["width"]), float(s["height"]), float(s["sigma"]))
        if s.get("width_dir") is not None:
            d = d.with_width_dir([float(x) for x in s["width_dir"]])
        g.add_segment(d)
    ports = []
    for p in doc["ports"]:
        _check_keys(p, ("positive", "negative", "name"), "port")
        ports.append(Port(int(p["positive"]), int(p["negative"]), p.get("name")))
    disc = doc["discretization"]
    if not isinstance(disc, dict) or len(disc) != 1:
        raise InvalidArgument("discretization must name exactly one variant")
    (kind, value), = disc.items()
    if kind == "uniform":
        dz = Discretization.uniform(int(value["nw"]), int(value["nh"]))
    elif kind == "per_segment":
        dz = Discretization.per_segment([Subdivision(int(v["nw"]), int(v["nh"])) for v in value])
    elif kind == "graded":
        dz = Discretization("graded", Grading(int(value["nw"]), int(value["nh"]), float(value["ratio"])))
    elif kind == "skin_depth":
        dz = Discretization.skin_depth_with(SkinDepthGrading(
            float(value["frequency_hz"]), float(value["target_skin_depths"]), float(value["ratio"]),
            int(value.get("max_per_axis", SkinDepthGrading.DEFAULT_MAX_PER_AXIS))))
    elif kind == "per_segment_graded":
        dz = Discretization.per_segment_graded(
            [AxisGrading(int(v["nw"]), int(v["nh"]), float(v["width_ratio"]), float(v["height_ratio"]))
             for v in value])
    else:
        raise InvalidArgument(f"unknown discretization variant `{kind}`")
    coupling = Coupling.all_pairs()
    if "coupling" in doc:
        c = doc["coupling"]
        _check_keys(c, ("segment_groups", "couples"), "coupling")
        couples = c.get("couples")
        coupling = Coupling(c.get("segment_groups", []),
                            None if couples is None else [tuple(x) for x in couples])
    return Deck(None, g, ports, dz, coupling, [float(f) for f in doc["frequencies_hz"]])


def problem_to_json(deck: Deck) -> str:
    import json

    def disc(d: Discretization):
        v = d.value
        if d.kind == "uniform":
            return {"uniform": {"nw": v.nw, "nh": v.nh}}
        if d.kind == "per_segment":
            return {"per_segment": [{"nw": s.nw, "nh": s.nh} for s in v]}
        if d.kind == "graded":
            return {"graded": {"nw": v.nw, "nh": v.nh, "ratio": v.ratio}}
        if d.kind == "skin_depth":
            return {"skin_depth": {"frequency_hz": v.frequency_hz, "target_skin_depths": v.target_skin_depths,
                                   "ratio": v.ratio, "max_per_axis": v.max_per_axis}}
        return {"per_segment_graded": [{"nw": a.nw, "nh": a.nh, "width_ratio": a.width_ratio,
                                        "height_ratio": a.height_ratio} for a in v]}

    doc = {
        "geometry": {
            "nodes": [{"x": n.x, "y": n.y, "z": n.z} for n in deck.geometry.nodes],
            "segments": [dict({"a": s.a, "b": s.b, "width": s.width, "height": s.height, "sigma": s.sigma},
                              **({"width_dir": list(s.width_dir)} if s.width_dir is not None else {}))
                         for s in deck.geometry.segments],
        },
        "ports": [dict({"positive": p.positive, "negative": p.negative},
                       **({"name": p.name} if p.name is not None else {})) for p in deck.ports],
        "discretization": disc(deck.discretization),
        "frequencies_hz": list(deck.frequencies),
    }
    if not deck.coupling.is_trivial():
        c = {}
        if deck.coupling.segment_groups:
            c["segment_groups"] = list(deck.coupling.segment_groups)
        if not deck.coupling.is_all_pairs():
            c["couples"] = [list(x) for x in (deck.coupling._couples or [])]
        doc["coupling"] = c
    return json.dumps(doc)
