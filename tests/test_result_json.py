"""SweepResult's schema-1 JSON (fasterhenry/src/result.rs:1-186), as the
reference's serde layout reads and writes it.

The layout checks restate tests/solve.rs:682-751
(result_serializes_to_json_with_provenance); the documented example of
result.rs:7-22 is read back as a fixture.
"""
import json
import math

import numpy as np
import pytest

from fasterhenry_amd import Counts, SweepResult, Timing, fixtures, model, solve
from fasterhenry_amd.errors import InvalidArgument

# result.rs:7-22, the module documentation's example document
DOC_EXAMPLE = """{
  "schema_version": 1,
  "ports": [{ "positive": 0, "negative": 1, "name": "trace" }],
  "frequencies_hz": [0.0, 1000000.0],
  "impedance_ohm": [
    [[{ "re": 0.0049, "im": 0.0 }]],
    [[{ "re": 0.0051, "im": 0.0437 }]]
  ],
  "provenance": {
    "fasterhenry_version": "0.1.0",
    "counts": { "nodes": 2, "segments": 1, "filaments": 10,
                "meshes": 10, "internal_meshes": 9, "ports": 1 },
    "timing": { "threads": 8, "inductance_s": 0.0004, "assembly_s": 0.0005,
                "solve_s": 0.0001, "total_s": 0.0006 }
  }
}"""

TOP_KEYS = {"schema_version", "ports", "frequencies_hz", "impedance_ohm", "provenance"}
COUNT_KEYS = {"nodes", "segments", "filaments", "meshes", "internal_meshes", "ports"}
TIMING_KEYS = {"threads", "inductance_s", "assembly_s", "solve_s", "total_s"}


def synthetic(p=3, k=3, timing=True):
    rng = np.random.default_rng(3)
    z = rng.standard_normal((k, p, p)) + 1j * rng.standard_normal((k, p, p))
    z[0] = z[0].real
    ports = [model.Port(0, 1, "A")] + [model.Port(2 * i, 2 * i + 1) for i in range(1, p)]
    counts = Counts(nodes=11, segments=8, filaments=41, meshes=30, internal_meshes=27, ports=p)
    t = Timing(8, 0.25, 0.5, 0.125, 0.625) if timing else None
    return SweepResult(ports, [0.0, 1e6, 1e9][:k], z, counts, t)


def check_layout(doc, result):
    assert set(doc) == TOP_KEYS
    assert doc["schema_version"] == 1
    prov = doc["provenance"]
    assert isinstance(prov["fasterhenry_version"], str)
    assert set(prov["counts"]) == COUNT_KEYS
    assert all(isinstance(v, int) for v in prov["counts"].values())
    if result.timing is not None:
        assert TIMING_KEYS <= set(prov["timing"])
        assert isinstance(prov["timing"]["threads"], int) and prov["timing"]["threads"] >= 1
        for key in TIMING_KEYS - {"threads"}:
            assert isinstance(prov["timing"][key], float)
    else:
        assert "timing" not in prov
    assert doc["frequencies_hz"] == [float(f) for f in result.frequencies_hz]
    mats = doc["impedance_ohm"]
    assert len(mats) == len(result.frequencies_hz)
    p = len(result.ports)
    for k, m in enumerate(mats):
        assert len(m) == p
        for i in range(p):
            assert len(m[i]) == p
            for j in range(p):
                assert set(m[i][j]) == {"re", "im"}
                assert m[i][j]["re"] == result.impedance_ohm[k][i, j].real
                assert m[i][j]["im"] == result.impedance_ohm[k][i, j].imag


def same(a, b):
    assert [(p.positive, p.negative, p.name) for p in a.ports] == \
        [(p.positive, p.negative, p.name) for p in b.ports]
    assert a.frequencies_hz == b.frequencies_hz
    assert a.impedance_ohm.tobytes() == b.impedance_ohm.tobytes()
    assert a.counts == b.counts
    assert (a.timing is None) == (b.timing is None)
    if a.timing is not None:
        assert a.timing.as_dict() == b.timing.as_dict()


def test_layout_and_exact_round_trip():
    r = synthetic()
    text = r.to_json()
    doc = json.loads(text)
    check_layout(doc, r)
    assert doc["ports"][0] == {"positive": 0, "negative": 1, "name": "A"}
    assert "name" not in doc["ports"][1]
    same(SweepResult.from_json(text), r)


def test_without_timing_drops_only_the_timing_block():
    r = synthetic()
    a = r.without_timing()
    assert a.timing is None and r.timing is not None
    doc = json.loads(a.to_json())
    check_layout(doc, a)
    assert "timing" not in doc["provenance"]
    assert a.to_json() == synthetic(timing=False).to_json()


def test_reads_the_documented_example():
    r = SweepResult.from_json(DOC_EXAMPLE)
    assert r.counts == Counts(2, 1, 10, 10, 9, 1)
    assert r.timing.threads == 8 and r.timing.total_s == 0.0006
    assert r.impedance_ohm.shape == (2, 1, 1)
    assert r.impedance_ohm[1, 0, 0] == complex(0.0051, 0.0437)
    assert r.ports[0].name == "trace"
    # written back, the reference's keys and values are unchanged
    doc, ref = json.loads(r.to_json()), json.loads(DOC_EXAMPLE)
    for key in ("schema_version", "ports", "frequencies_hz", "impedance_ohm"):
        assert doc[key] == ref[key]
    assert doc["provenance"]["counts"] == ref["provenance"]["counts"]
    for key in TIMING_KEYS:
        assert doc["provenance"]["timing"][key] == ref["provenance"]["timing"][key]


def test_rejects_a_non_square_matrix_and_non_finite_values():
    doc = json.loads(synthetic().to_json())
    doc["impedance_ohm"][0][0].pop()
    with pytest.raises(InvalidArgument):
        SweepResult.from_json(json.dumps(doc))
    bad = synthetic()
    bad.impedance_ohm[1, 0, 0] = complex(math.nan, 0.0)
    with pytest.raises(ValueError):
        bad.to_json()


@pytest.mark.gpu
def test_solved_result_serializes_with_provenance():
    """tests/solve.rs:682-751 on the coupled structure: counts, timing keys,
    frequencies, port names, {re, im} entries, an exact round trip, and two
    runs identical without timing."""
    g, ports, d = fixtures.coupled_structure()
    freqs = [0.0, 1e6, 1e9]
    result = solve(g, ports, d, freqs)
    doc = json.loads(result.to_json())
    check_layout(doc, result)
    counts = doc["provenance"]["counts"]
    assert counts["segments"] == 8 and counts["nodes"] == 11 and counts["ports"] == 3
    assert counts["filaments"] == 4 * 3 + 3 * 3 + 3 + 1 + 4 * 4
    assert counts["meshes"] == counts["internal_meshes"] + 3
    assert doc["frequencies_hz"] == [0.0, 1e6, 1e9]
    assert doc["ports"][0]["name"] == "A" and doc["ports"][1]["positive"] == 4
    same(SweepResult.from_json(result.to_json()), result)
    again = solve(g, ports, d, freqs)
    assert result.without_timing().to_json() == again.without_timing().to_json()
