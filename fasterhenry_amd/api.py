"""The library surface: solve(), MeshSystem, SweepResult, kernel-level calls.

Mirrors fasterhenry 0.2.0:
* ``solve`` -- fasterhenry/src/solve.rs:931-938
* ``MeshSystem.assemble/impedance/sweep/counts`` -- solve.rs:643-895
* ``SweepResult`` (schema 1) -- result.rs:47-94
* ``partial_inductance_matrix`` -- inductance/batch.rs:246-253
* ``discretize_geometry`` -- solve.rs:947-975
* ``IterativeSystem`` -- iterative.rs:155-536; ``PfftOperator`` --
  pfft/mod.rs:270-528; ``gmres_dense`` -- gmres.rs:91-206

Every call runs on the GPU through libfasterhenry_gpu.so; there is no CPU
fallback.  ``precision`` selects the factorization arithmetic: "fp64"
(plain) or "fp32_refined" (FP32 matrix-core factorization + FP64
refinement).
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import native
from .coupling import Coupling, PairMask
from .errors import InvalidArgument, SolveError
from .model import (DENSE_PATH_MAX_FILAMENTS, FILAMENT_DTYPE, Discretization, Geometry,
                    PackedProblem, Port, Solver, SolverChoice, pack)

VERSION = "0.1.0"

_PRECISIONS = {"fp64": native.FH_FP64_PLAIN, "fp32_refined": native.FH_FP32_REFINED}


def _precision_code(precision: str) -> int:
    try:
        return _PRECISIONS[precision]
    except KeyError:
        raise InvalidArgument(f"precision must be one of {sorted(_PRECISIONS)}, got {precision!r}")


@dataclass
class Counts:
    """result.rs:110-125"""

    nodes: int
    segments: int
    filaments: int
    meshes: int
    internal_meshes: int
    ports: int


def host_threads() -> int:
    """Host threads the library's parallel host work may use (the pFFT
    near-pair search caps its pool at 16): the ``threads`` of ``Timing``."""
    return max(1, min(16, os.cpu_count() or 1))


@dataclass
class Timing:
    """result.rs:129-140.  ``device`` is an extra key naming the GPU; a
    reader of the reference layout (serde, which ignores unknown keys here)
    sees the reference's five fields."""

    threads: int
    inductance_s: float
    assembly_s: float
    solve_s: float
    total_s: float
    device: str = "gfx950"

    _FIELDS = ("threads", "inductance_s", "assembly_s", "solve_s", "total_s")

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k in self._FIELDS}
        d["threads"] = int(d["threads"])
        d["device"] = self.device
        return d


@dataclass
class SweepResult:
    """Port impedance matrices over frequencies, with provenance (result.rs:47-94)."""

    ports: List[Port]
    frequencies_hz: List[float]
    impedance_ohm: np.ndarray  # (K, p, p) complex128
    counts: Counts
    timing: Optional[Timing] = None
    schema_version: int = 1
    extra: Dict[str, float] = field(default_factory=dict)

    SCHEMA_VERSION = 1

    def resistance(self, index: int) -> np.ndarray:
        return self.impedance_ohm[index].real.copy()

    def inductance(self, index: int) -> Optional[np.ndarray]:
        omega = 2.0 * math.pi * self.frequencies_hz[index]
        return None if omega == 0.0 else self.impedance_ohm[index].imag / omega

    def without_timing(self) -> "SweepResult":
        """The deterministic part only (result.rs SweepResult::without_timing)."""
        return SweepResult(list(self.ports), list(self.frequencies_hz), self.impedance_ohm.copy(),
                           self.counts, None, self.schema_version, dict(self.extra))

    def to_json(self) -> str:
        """The reference's JSON layout (result.rs:1-37): ``impedance_ohm[k][i][j]``
        as ``{"re", "im"}`` objects, rows following ``ports``; timing only when
        present.  Every number must be finite (result.rs:32)."""
        doc = {
            "schema_version": self.schema_version,
            "ports": [{"positive": int(p.positive), "negative": int(p.negative),
                       **({"name": p.name} if p.name is not None else {})} for p in self.ports],
            "frequencies_hz": [float(f) for f in self.frequencies_hz],
            "impedance_ohm": [[[{"re": float(z.real), "im": float(z.imag)} for z in row]
                               for row in m] for m in self.impedance_ohm],
            "provenance": {
                "fasterhenry_version": f"fasterhenry_amd {VERSION}",
                "counts": {k: int(v) for k, v in self.counts.__dict__.items()},
                **({"timing": self.timing.as_dict()} if self.timing else {}),
            },
        }
        return json.dumps(doc, allow_nan=False)

    @classmethod
    def from_json(cls, text: str) -> "SweepResult":
        """Read the schema-1 layout, as written here or by the reference."""
        doc = json.loads(text)
        if doc.get("schema_version") != cls.SCHEMA_VERSION:
            raise InvalidArgument(f"unsupported schema_version {doc.get('schema_version')!r}")
        ports = [Port(int(p["positive"]), int(p["negative"]), p.get("name")) for p in doc["ports"]]
        freqs = [float(f) for f in doc["frequencies_hz"]]
        mats = []
        for rows in doc["impedance_ohm"]:
            if any(len(r) != len(rows) for r in rows):
                raise InvalidArgument("impedance matrix is not square")
            mats.append([[complex(e["re"], e["im"]) for e in r] for r in rows])
        z = np.array(mats, dtype=np.complex128).reshape(len(mats), len(ports), len(ports))
        prov = doc["provenance"]
        counts = Counts(**{k: int(prov["counts"][k]) for k in Counts.__dataclass_fields__})
        t = prov.get("timing")
        timing = None
        if t is not None:
            timing = Timing(int(t["threads"]), float(t["inductance_s"]), float(t["assembly_s"]),
                            float(t["solve_s"]), float(t["total_s"]), str(t.get("device", "")))
        return cls(ports, freqs, z, counts, timing, int(doc["schema_version"]))

    def to_zc_mat(self) -> bytes:
        """MAT v4 bytes of ``freqs`` and ``Zc_1 .. Zc_K`` (fasterhenry-cli/src/mat.rs:31-95)."""
        return zc_mat(self.frequencies_hz, self.impedance_ohm)

    def write_zc_mat(self, path: str) -> None:
        with open(path, "wb") as fh:
            fh.write(self.to_zc_mat())

    def to_spice(self, index: int, name: str) -> str:
        """The SPICE subcircuit stamp at frequency ``index``
        (fasterhenry-cli/src/spice.rs:64-126)."""
        return spice_subckt(self.frequencies_hz, self.impedance_ohm, index, name)


def _z_array(frequencies, z) -> tuple:
    f = np.ascontiguousarray(frequencies, dtype=np.float64).reshape(-1)
    zz = np.ascontiguousarray(z, dtype=np.complex128)
    if zz.ndim != 3 or zz.shape[0] != len(f) or zz.shape[1] != zz.shape[2]:
        raise InvalidArgument(f"z must be (K, p, p) for K = {len(f)} frequencies, got {zz.shape}")
    return f, zz


def zc_mat(frequencies, z) -> bytes:
    """write_zc_mat (fasterhenry-cli/src/mat.rs:31-95) of a (K, p, p) sweep."""
    f, zz = _z_array(frequencies, z)
    L = native.lib()
    size = C.c_size_t()
    native.check(L.fh_zc_mat(native.ptr(f), len(f), native.ptr(zz), zz.shape[1], None, 0, C.byref(size)))
    out = (C.c_uint8 * max(size.value, 1))()
    native.check(L.fh_zc_mat(native.ptr(f), len(f), native.ptr(zz), zz.shape[1], out, size.value,
                             C.byref(size)))
    return bytes(out[:size.value])


def spice_subckt(frequencies, z, index: int, name: str) -> str:
    """write_spice_subckt (fasterhenry-cli/src/spice.rs:64-126)."""
    f, zz = _z_array(frequencies, z)
    L = native.lib()
    length = C.c_size_t()
    nm = name.encode()
    native.check(L.fh_spice_subckt(native.ptr(f), len(f), native.ptr(zz), zz.shape[1], index, nm, None, 0,
                                   C.byref(length)))
    buf = C.create_string_buffer(length.value + 1)
    native.check(L.fh_spice_subckt(native.ptr(f), len(f), native.ptr(zz), zz.shape[1], index, nm, buf,
                                   length.value + 1, C.byref(length)))
    return buf.value.decode()


def subckt_name_from_path(path: str) -> str:
    """subckt_name_from_path (fasterhenry-cli/src/spice.rs:37-60)."""
    L = native.lib()
    length = C.c_size_t()
    raw = path.encode("utf-8", "surrogateescape")
    native.check(L.fh_subckt_name_from_path(raw, None, 0, C.byref(length)))
    buf = C.create_string_buffer(length.value + 1)
    native.check(L.fh_subckt_name_from_path(raw, buf, length.value + 1, C.byref(length)))
    return buf.value.decode()


@dataclass
class FilamentCurrent:
    """Current in one filament (currents.rs:54-71)."""

    filament: int
    segment: int
    conductor: int
    location: List[float]
    current: complex
    current_density: float
    power_w: float


@dataclass
class SegmentCurrent:
    """Total current through, and power in, one segment (currents.rs:74-84)."""

    segment: int
    conductor: int
    current: complex
    power_w: float


@dataclass
class ConductorPeak:
    """The hottest filament of one conductor (currents.rs:87-97)."""

    conductor: int
    peak_current_density: float
    filament: int
    location: List[float]


@dataclass
class PowerBalance:
    """Dissipated versus delivered power (currents.rs:100-108)."""

    dissipated_w: float
    input_w: float
    relative_residual: float


@dataclass
class CurrentSolution:
    """All currents of one frequency and excitation (currents.rs:111-124):
    peak phasors in amperes, |J| in A/m^2, time-average power in watts."""

    frequency_hz: float
    excitation: List[complex]
    filaments: List[FilamentCurrent]
    segments: List[SegmentCurrent]
    conductors: List[ConductorPeak]
    power_balance: PowerBalance

    def to_dict(self) -> dict:
        def c(z):
            return {"re": float(z.real), "im": float(z.imag)}
        return {
            "frequency_hz": self.frequency_hz,
            "excitation": [c(z) for z in self.excitation],
            "filaments": [dict(filament=f.filament, segment=f.segment, conductor=f.conductor,
                               location=f.location, current=c(f.current),
                               current_density=f.current_density, power_w=f.power_w)
                          for f in self.filaments],
            "segments": [dict(segment=g.segment, conductor=g.conductor, current=c(g.current),
                              power_w=g.power_w) for g in self.segments],
            "conductors": [dict(conductor=k.conductor, peak_current_density=k.peak_current_density,
                                filament=k.filament, location=k.location) for k in self.conductors],
            "power_balance": dict(dissipated_w=self.power_balance.dissipated_w,
                                  input_w=self.power_balance.input_w,
                                  relative_residual=self.power_balance.relative_residual),
        }


def unit_excitation(ports: int) -> List[complex]:
    """1 A into port 1, every other port open (currents.rs:126-132)."""
    return [complex(1.0 if i == 0 else 0.0, 0.0) for i in range(ports)]


def _currents(handle, counts, packed: PackedProblem, filaments: np.ndarray, frequency: float,
              excitation) -> CurrentSolution:
    ex = np.ascontiguousarray(np.asarray(excitation, dtype=np.complex128).reshape(-1))
    n, nseg, p = counts.filaments, counts.segments, counts.ports
    cur = np.zeros(n, dtype=np.complex128)
    dens = np.zeros(n)
    power = np.zeros(n)
    scur = np.zeros(nseg, dtype=np.complex128)
    spow = np.zeros(nseg)
    cond = np.zeros(nseg, dtype=np.int32)
    peaks = np.zeros(max(nseg, 1), dtype=np.int64)
    volt = np.zeros(max(p, 1), dtype=np.complex128)
    ncond = C.c_size_t()
    bal = native.PowerBalance()
    native.check(native.lib().fh_currents(
        handle, float(frequency), native.ptr(ex) if len(ex) else None, len(ex), native.ptr(cur),
        native.ptr(dens), native.ptr(power), native.ptr(scur), native.ptr(spow), native.ptr(cond),
        C.byref(ncond), native.ptr(peaks), native.ptr(volt), C.byref(bal)))
    # Filament::center (filament.rs:91-93)
    centre = (np.asarray(filaments["start"]) + np.asarray(filaments["end"])) * 0.5
    # filaments come segment by segment
    per = packed.grids["nw"].astype(np.int64) * packed.grids["nh"].astype(np.int64)
    seg_of = np.repeat(np.arange(nseg), per)
    fils = []
    for k in range(n):
        q = int(seg_of[k])
        fils.append(FilamentCurrent(k, q, int(cond[q]), [float(x) for x in centre[k]], complex(cur[k]),
                                    float(dens[k]), float(power[k])))
    segs = [SegmentCurrent(q, int(cond[q]), complex(scur[q]), float(spow[q])) for q in range(nseg)]
    conds = [ConductorPeak(c, float(dens[peaks[c]]), int(peaks[c]),
                           [float(x) for x in centre[peaks[c]]]) for c in range(ncond.value)]
    return CurrentSolution(float(frequency), [complex(z) for z in ex], fils, segs, conds,
                           PowerBalance(bal.dissipated_w, bal.input_w, bal.relative_residual))


def _counts(handle) -> Counts:
    c = (C.c_size_t * 6)()
    native.check(native.lib().fh_system_counts(handle, c))
    return Counts(*[int(x) for x in c])


class MeshSystem:
    """An assembled mesh system resident in GPU memory (solve.rs:643-895)."""

    def __init__(self, handle, packed: PackedProblem, ports: Sequence[Port], precision: str):
        self._h = handle
        self._packed = packed
        self.ports = list(ports)
        self.precision = precision
        self._counts = _counts(handle)
        self.truncation_warnings: List = []

    @classmethod
    def assemble(cls, geometry: Geometry, ports: Sequence[Port], discretization: Discretization,
                 precision: str = "fp64", solver: SolverChoice = SolverChoice.DENSE) -> "MeshSystem":
        packed = pack(geometry, ports, discretization)
        return cls.assemble_packed(packed, ports, precision, solver)

    @classmethod
    def assemble_with_coupling(cls, geometry: Geometry, ports: Sequence[Port],
                               discretization: Discretization, coupling: Coupling,
                               precision: str = "fp64",
                               solver: SolverChoice = SolverChoice.DENSE) -> "MeshSystem":
        """MeshSystem::assemble_with_coupling (solve.rs:703-766): the mutual
        inductance of uncoupled groups is never evaluated.  The coupling is
        validated (and its warnings computed) before anything is assembled."""
        warnings = coupling.truncation_warnings(geometry)
        packed = pack(geometry, ports, discretization)
        if coupling.is_all_pairs():
            system = cls.assemble_packed(packed, ports, precision, solver)
        else:
            group, groups, table = coupling.resolved(geometry.segment_count())
            cm = native.CouplingMask(native.ptr(group) if len(group) else None, len(group), groups,
                                     native.ptr(table))
            L = native.lib()
            h = C.c_void_p()
            prob = native.problem_struct(packed)
            native.check(L.fh_system_create_coupled(C.byref(prob), solver.value, _precision_code(precision),
                                                    C.byref(cm), C.byref(h)))
            system = cls(h, packed, ports, precision)
        system.truncation_warnings = warnings
        return system

    @classmethod
    def assemble_packed(cls, packed: PackedProblem, ports: Sequence[Port], precision: str = "fp64",
                        solver: SolverChoice = SolverChoice.DENSE) -> "MeshSystem":
        L = native.lib()
        h = C.c_void_p()
        prob = native.problem_struct(packed)
        native.check(L.fh_system_create(C.byref(prob), solver.value, _precision_code(precision),
                                        C.byref(h)))
        return cls(h, packed, ports, precision)

    def close(self) -> None:
        if getattr(self, "_h", None):
            native.lib().fh_system_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, precision: str) -> None:
        native.check(native.lib().fh_system_set_precision(self._h, _precision_code(precision)))
        self.precision = precision

    def set_profiling(self, enable: bool) -> None:
        native.check(native.lib().fh_system_set_profiling(self._h, 1 if enable else 0))

    def set_graphs(self, enable: bool) -> None:
        """Graph replay of repeated refined sweeps (fh_system_set_graphs)."""
        native.check(native.lib().fh_system_set_graphs(self._h, 1 if enable else 0))

    def schedule(self) -> dict:
        """The sweep's schedule (fh_system_get_schedule)."""
        sc = native.Schedule()
        native.check(native.lib().fh_system_get_schedule(self._h, C.byref(sc)))
        return sc.as_dict()

    def set_schedule(self, outer_block: int, sweep_parts: int) -> None:
        """Fix the sweep's schedule (fh_system_set_schedule)."""
        sc = native.Schedule(int(outer_block), int(sweep_parts))
        native.check(native.lib().fh_system_set_schedule(self._h, C.byref(sc)))

    def autotune(self, frequencies) -> tuple:
        """Time the candidate schedules on this system's sweep of
        ``frequencies``, keep the fastest (fh_system_autotune).  Returns
        (schedule dict, its sweep time in ms)."""
        f = np.ascontiguousarray(frequencies, dtype=np.float64)
        sc = native.Schedule()
        ms = C.c_double()
        native.check(native.lib().fh_system_autotune(self._h, native.ptr(f), len(f), C.byref(sc),
                                                     C.byref(ms)))
        return sc.as_dict(), ms.value

    def set_refinement(self, stages: int) -> None:
        """The refinement stages the refined form may use (a mask of
        native.FH_REFINE_*; fh_system_set_refinement)."""
        native.check(native.lib().fh_system_set_refinement(self._h, int(stages)))

    def set_sweep_parts(self, parts: int) -> None:
        """Refined dense sweeps: factorize a chunk of frequencies as 1 to 4
        groups on as many streams (default 2; a group has at least 4
        frequencies)."""
        native.check(native.lib().fh_system_set_sweep_parts(self._h, int(parts)))

    def counts(self) -> Counts:
        return self._counts

    def sweep_raw(self, frequencies) -> tuple:
        f = np.ascontiguousarray(frequencies, dtype=np.float64)
        p = self._counts.ports
        z = np.zeros((len(f), p, p), dtype=np.complex128)
        info = native.SweepInfo()
        native.check(native.lib().fh_sweep(self._h, native.ptr(f), len(f), native.ptr(z),
                                           C.byref(info)))
        return z, info

    def sweep(self, frequencies) -> SweepResult:
        """Z at every frequency (solve.rs:881-894)."""
        z, info = self.sweep_raw(frequencies)
        timing = Timing(host_threads(), info.inductance_s, info.assembly_s, info.solve_s,
                        info.assembly_s + info.solve_s)
        return SweepResult(self.ports, [float(x) for x in np.asarray(frequencies, dtype=float)], z,
                           self._counts, timing,
                           extra={"refine_steps_max": info.refine_steps_max,
                                  "refine_change_max": info.refine_change_max,
                                  "factor_arith": info.factor_arith,
                                  "refine_stage": info.refine_stage})

    def impedance(self, frequency: float) -> np.ndarray:
        """Z at one frequency (solve.rs:806-830)."""
        return self.sweep_raw([frequency])[0][0]

    def filaments(self) -> np.ndarray:
        out = np.zeros(self._counts.filaments, dtype=FILAMENT_DTYPE)
        native.check(native.lib().fh_system_filaments(self._h, native.ptr(out)))
        return out

    def inductance_matrix(self) -> np.ndarray:
        n = self._counts.filaments
        out = np.zeros((n, n))
        native.check(native.lib().fh_system_inductance(self._h, native.ptr(out)))
        return out

    def currents(self, frequency: float, excitation) -> CurrentSolution:
        """Filament currents, |J| and power for port-current phasors
        ``excitation`` (MeshSystem::currents, solve.rs:835-868)."""
        if getattr(self, "_fil", None) is None:
            self._fil = self.filaments()
        return _currents(self._h, self._counts, self._packed, self._fil, frequency, excitation)

    def grad(self, frequencies, d_nodes, d_widths=None, d_heights=None) -> tuple:
        """Z and its derivatives along n_params geometry directions (fh_grad;
        the reference has no gradients).  d_nodes: (n_params, n_nodes, 3) node
        velocities; d_widths / d_heights: (n_params, n_segments) or None.
        Returns (z (K, p, p), dR_dp (K, n_params, p, p), dL_dp (K, n_params,
        p, p)) with dR = Re dZ/dp and dL = Im dZ/dp / w (0 at f = 0)."""
        f = np.ascontiguousarray(frequencies, dtype=np.float64)
        K, p = len(f), self._counts.ports
        arrays, pmap = _param_map(self._packed, d_nodes, d_widths, d_heights)
        P = pmap.n_params
        z = np.zeros((K, p, p), dtype=np.complex128)
        dr = np.zeros((K, P, p, p))
        dl = np.zeros((K, P, p, p))
        native.check(native.lib().fh_grad(self._h, C.byref(pmap), native.ptr(f), K, native.ptr(z),
                                          native.ptr(dr), native.ptr(dl)))
        del arrays
        return z, dr, dl

    def blocks(self) -> Dict[str, np.ndarray]:
        m, p = self._counts.internal_meshes, self._counts.ports
        out = [np.zeros((m, m)), np.zeros((m, m)), np.zeros((m, p)), np.zeros((m, p)),
               np.zeros((p, p)), np.zeros((p, p))]
        native.check(native.lib().fh_system_blocks(self._h, *(native.ptr(x) for x in out)))
        return dict(zip(["r_ee", "l_ee", "r_ep", "l_ep", "r_pp", "l_pp"], out))


def pfft_params(spacing: Optional[float] = None, cells_per_filament: float = 8.0,
                near_field_radius: float = 8.0, interpolation_order: int = 5,
                fft_precision: str = "fp64") -> native.PfftParams:
    """PfftParams (pfft/mod.rs:183-231): GridSpacing::Fixed(spacing) when
    spacing is given, else GridSpacing::Auto{cells_per_filament}.
    fft_precision "fp32" runs the far-field FFT convolution in FP32."""
    if fft_precision not in ("fp64", "fp32"):
        raise InvalidArgument(f"fft_precision must be 'fp64' or 'fp32', got {fft_precision!r}")
    p = native.PfftParams()
    native.lib().fh_pfft_default_params(C.byref(p))
    p.auto_grid = 0 if spacing is not None else 1
    p.spacing = float(spacing or 0.0)
    p.cells_per_filament = float(cells_per_filament)
    p.near_field_radius = float(near_field_radius)
    p.interpolation_order = int(interpolation_order)
    p.fft_precision = 1 if fft_precision == "fp32" else 0
    return p


def iterative_params(pfft: Optional[native.PfftParams] = None, tolerance: float = 1e-10,
                     restart: int = 100, max_iterations: int = 5000) -> native.IterativeParams:
    """IterativeParams (iterative.rs:98-105) with GmresParams (gmres.rs:29-60)."""
    p = native.IterativeParams()
    native.lib().fh_iterative_default_params(C.byref(p))
    if pfft is not None:
        p.pfft = pfft
    p.gmres.tolerance = float(tolerance)
    p.gmres.restart = int(restart)
    p.gmres.max_iterations = int(max_iterations)
    return p


class IterativeSystem:
    """The matrix-free mesh system (iterative.rs:155-536): pFFT operator,
    port columns and Jacobi diagonals resident on the GPU; one GMRES solve
    per port and frequency."""

    def __init__(self, handle, packed: PackedProblem, ports: Sequence[Port]):
        self._h = handle
        self._packed = packed
        self.ports = list(ports)
        self._counts = _counts(handle)

    @classmethod
    def assemble(cls, geometry: Geometry, ports: Sequence[Port], discretization: Discretization,
                 params: Optional[native.IterativeParams] = None) -> "IterativeSystem":
        return cls.assemble_packed(pack(geometry, ports, discretization), ports, params)

    @classmethod
    def assemble_packed(cls, packed: PackedProblem, ports: Sequence[Port],
                        params: Optional[native.IterativeParams] = None) -> "IterativeSystem":
        L = native.lib()
        h = C.c_void_p()
        prob = native.problem_struct(packed)
        params = params if params is not None else iterative_params()
        native.check(L.fh_system_create_iterative(C.byref(prob), C.byref(params), C.byref(h)))
        return cls(h, packed, ports)

    def close(self) -> None:
        if getattr(self, "_h", None):
            native.lib().fh_system_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_profiling(self, enable: bool) -> None:
        native.check(native.lib().fh_system_set_profiling(self._h, 1 if enable else 0))

    def counts(self) -> Counts:
        return self._counts

    def operator_stats(self) -> dict:
        """IterativeSystem::operator_stats (iterative.rs:255-258)."""
        st = native.PfftStats()
        native.check(native.lib().fh_system_operator_stats(self._h, C.byref(st)))
        return st.as_dict()

    def filaments(self) -> np.ndarray:
        out = np.zeros(self._counts.filaments, dtype=FILAMENT_DTYPE)
        native.check(native.lib().fh_system_filaments(self._h, native.ptr(out)))
        return out

    sweep_raw = MeshSystem.sweep_raw

    def sweep(self, frequencies) -> SweepResult:
        """Z at every frequency (iterative.rs:387-404)."""
        z, info = self.sweep_raw(frequencies)
        timing = Timing(host_threads(), info.inductance_s, info.assembly_s, info.solve_s,
                        info.assembly_s + info.solve_s)
        return SweepResult(self.ports, [float(x) for x in np.asarray(frequencies, dtype=float)], z,
                           self._counts, timing,
                           extra={"gmres_iterations_max": info.gmres_iterations_max,
                                  "gmres_iterations": info.gmres_iterations,
                                  "gmres_residual_max": info.gmres_residual_max})

    def impedance(self, frequency: float) -> np.ndarray:
        """Z at one frequency (iterative.rs:270-272)."""
        return self.sweep_raw([frequency])[0][0]

    def currents(self, frequency: float, excitation) -> CurrentSolution:
        """Filament currents, |J| and power for port-current phasors
        ``excitation`` (IterativeSystem::currents, iterative.rs:329-380)."""
        if getattr(self, "_fil", None) is None:
            self._fil = self.filaments()
        return _currents(self._h, self._counts, self._packed, self._fil, frequency, excitation)


class PfftOperator:
    """The precorrected-FFT partial-inductance operator (pfft/mod.rs:270-528)
    of a FILAMENT_DTYPE array, resident on the GPU."""

    def __init__(self, filaments: np.ndarray, params: Optional[native.PfftParams] = None):
        self._f = np.ascontiguousarray(filaments, dtype=FILAMENT_DTYPE)
        self.n = len(self._f)
        self.params = params if params is not None else pfft_params()
        self._h = C.c_void_p()
        native.check(native.lib().fh_pfft_create(native.ptr(self._f) if self.n else None, self.n,
                                                 C.byref(self.params), C.byref(self._h)))

    def close(self) -> None:
        if getattr(self, "_h", None):
            native.lib().fh_pfft_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def apply(self, x) -> np.ndarray:
        """y = L x (pfft/mod.rs:482-528) for one vector (n,) or a batch (V, n)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        nvec = 1 if x.ndim == 1 else x.shape[0]
        if x.shape[-1] != self.n:
            raise InvalidArgument(f"vector length {x.shape[-1]} != {self.n} filaments")
        native.check(native.lib().fh_pfft_apply(self._h, native.ptr(x), native.ptr(y), nvec))
        return y

    def stats(self) -> dict:
        """PfftStats (pfft/mod.rs:255-270)."""
        st = native.PfftStats()
        native.check(native.lib().fh_pfft_get_stats(self._h, C.byref(st)))
        return st.as_dict()

    def data(self) -> dict:
        """The projection and near-field CSR arrays, copied from the device."""
        st = self.stats()
        pe, ne = st["projection_entries"], st["near_entries"]
        proj_off = np.zeros(self.n + 1, dtype=np.int64)
        proj_idx = np.zeros(max(pe, 1), dtype=np.uint32)
        proj_w = np.zeros(max(pe, 1))
        near_off = np.zeros(self.n + 1, dtype=np.int64)
        near_col = np.zeros(max(ne, 1), dtype=np.uint32)
        near_val = np.zeros(max(ne, 1))
        native.check(native.lib().fh_pfft_data(self._h, *(native.ptr(a) for a in (
            proj_off, proj_idx, proj_w, near_off, near_col, near_val))))
        return dict(proj_off=proj_off, proj_idx=proj_idx[:pe], proj_w=proj_w[:pe],
                    near_off=near_off, near_col=near_col[:ne], near_val=near_val[:ne], **st)


def gmres_dense(a: np.ndarray, b: np.ndarray, x0=None, tolerance: float = 1e-10,
                restart: int = 100, max_iterations: int = 5000,
                complex_arith: bool = True) -> tuple:
    """gmres (gmres.rs:91-206) on the GPU for a dense matrix, no
    preconditioner.  Returns (x, iterations, relative_residual, converged)."""
    a = np.asarray(a, dtype=np.complex128)
    n = a.shape[0]
    af = np.asfortranarray(a)
    bb = np.ascontiguousarray(b, dtype=np.complex128)
    x = np.zeros(n, dtype=np.complex128) if x0 is None else np.array(x0, dtype=np.complex128)
    prm = native.GmresParams(float(tolerance), int(restart), int(max_iterations))
    it, rr, cv = C.c_uint64(), C.c_double(), C.c_int32()
    native.check(native.lib().fh_gmres_dense(n, af.ctypes.data_as(C.c_void_p), native.ptr(bb),
                                             native.ptr(x), C.byref(prm),
                                             1 if complex_arith else 0, C.byref(it),
                                             C.byref(rr), C.byref(cv)))
    return x, it.value, rr.value, bool(cv.value)


def solve(geometry: Geometry, ports: Sequence[Port], discretization: Discretization,
          frequencies, solver: SolverChoice = SolverChoice.AUTO,
          precision: str = "fp64") -> SweepResult:
    """Assemble and sweep (solve.rs:931-938).  SolverChoice.AUTO resolves
    against DENSE_PATH_MAX_FILAMENTS: the dense GPU LU up to it, the pFFT +
    GMRES path (default IterativeParams) above."""
    n = discretization.filament_count(geometry)
    if solver.resolve(n) is Solver.DENSE:
        system = MeshSystem.assemble(geometry, ports, discretization, precision, SolverChoice.DENSE)
    else:
        system = IterativeSystem.assemble(geometry, ports, discretization)
    try:
        return system.sweep(frequencies)
    finally:
        system.close()


def solve_batch(problems, frequencies, solver: SolverChoice = SolverChoice.DENSE,
                precision: str = "fp32_refined", concurrency: int = 0) -> tuple:
    """solve() for every (geometry, ports, discretization) of ``problems``
    over the same frequencies, scheduled together on the device
    (fh_solve_batch).  Returns (list of (K, p, p) impedance arrays, list of
    per-problem sweep-info dicts)."""
    packed = [pack(g, ports, d) for (g, ports, d) in problems]
    structs = (native.Problem * len(packed))(*[native.problem_struct(pk) for pk in packed])
    f = np.ascontiguousarray(frequencies, dtype=np.float64)
    K = len(f)
    sizes = [len(pk.ports) for pk in packed]
    z = np.zeros(sum(K * p * p for p in sizes), dtype=np.complex128)
    infos = (native.SweepInfo * len(packed))()
    failed = C.c_size_t()
    native.check(native.lib().fh_solve_batch(structs, len(packed), solver.value, _precision_code(precision),
                                             native.ptr(f), K, native.ptr(z), infos, int(concurrency),
                                             C.byref(failed)))
    out, o = [], 0
    for p in sizes:
        out.append(z[o:o + K * p * p].reshape(K, p, p))
        o += K * p * p
    return out, [i.as_dict() for i in infos]


def tune_save(path: str) -> None:
    """Write the tuning cache to a JSON file (fh_tune_save)."""
    native.check(native.lib().fh_tune_save(str(path).encode()))


def tune_load(path: str) -> None:
    """Merge a tuning cache file into this process's cache (fh_tune_load)."""
    native.check(native.lib().fh_tune_load(str(path).encode()))


def discretize_geometry(geometry: Geometry, discretization: Discretization) -> np.ndarray:
    """All filaments, segment by segment (solve.rs:947-975), from the native host code."""
    packed = pack(geometry, [], discretization)
    return discretize_packed(packed)


def discretize_packed(packed: PackedProblem) -> np.ndarray:
    L = native.lib()
    prob = native.problem_struct(packed)
    n = C.c_size_t()
    native.check(L.fh_discretize(C.byref(prob), None, 0, C.byref(n)))
    out = np.zeros(n.value, dtype=FILAMENT_DTYPE)
    native.check(L.fh_discretize(C.byref(prob), native.ptr(out), n.value, C.byref(n)))
    return out


def _param_map(packed: PackedProblem, d_nodes, d_widths, d_heights) -> tuple:
    nn, ns = len(packed.nodes), len(packed.segments)
    dn = np.ascontiguousarray(d_nodes, dtype=np.float64)
    if dn.ndim == 2:
        dn = dn[None]
    P = dn.shape[0]
    if dn.shape != (P, nn, 3):
        raise InvalidArgument(f"d_nodes must be (n_params, {nn}, 3), got {dn.shape}")
    out = [dn]
    ptrs = []
    for d in (d_widths, d_heights):
        if d is None:
            ptrs.append(None)
            continue
        a = np.ascontiguousarray(d, dtype=np.float64).reshape(P, -1)
        if a.shape != (P, ns):
            raise InvalidArgument(f"segment derivatives must be (n_params, {ns}), got {a.shape}")
        out.append(a)
        ptrs.append(native.ptr(a))
    pmap = native.ParamMap(P, native.ptr(dn), ptrs[0], ptrs[1])
    return out, pmap


def discretize_tangent(packed: PackedProblem, d_nodes, d_widths=None, d_heights=None) -> np.ndarray:
    """d(filaments)/dp along one geometry direction (fh_discretize_tangent):
    d_nodes (n_nodes, 3), d_widths / d_heights (n_segments,) or None.  The
    filament counts and grading ratios are held at the problem's."""
    L = native.lib()
    prob = native.problem_struct(packed)
    n = C.c_size_t()
    native.check(L.fh_discretize(C.byref(prob), None, 0, C.byref(n)))
    arrays, pmap = _param_map(packed, np.asarray(d_nodes, dtype=np.float64)[None],
                              None if d_widths is None else np.asarray(d_widths)[None],
                              None if d_heights is None else np.asarray(d_heights)[None])
    out = np.zeros(n.value, dtype=FILAMENT_DTYPE)
    native.check(L.fh_discretize_tangent(C.byref(prob), pmap.d_nodes, pmap.d_widths, pmap.d_heights,
                                         native.ptr(out), n.value))
    del arrays
    return out


def partial_inductance_derivative(filaments: np.ndarray, tangents: np.ndarray, pairs) -> tuple:
    """Nominal partial inductances and their derivatives along filament
    tangents for a list of pairs (fh_partial_inductance_derivative).
    tangents: (n_dirs, n) filament records; pairs: (count, 3) rows of
    (f, g, direction).  Returns (values, derivatives)."""
    f = np.ascontiguousarray(filaments, dtype=FILAMENT_DTYPE)
    t = np.ascontiguousarray(tangents, dtype=FILAMENT_DTYPE)
    if t.ndim == 1:
        t = t[None]
    if t.shape[1] != len(f):
        raise InvalidArgument("tangents must be (n_dirs, n_filaments)")
    pr = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 3)
    v = np.zeros(len(pr))
    d = np.zeros(len(pr))
    native.check(native.lib().fh_partial_inductance_derivative(
        native.ptr(f), len(f), native.ptr(t), t.shape[0], native.ptr(pr), len(pr), native.ptr(v),
        native.ptr(d)))
    return v, d


def partial_inductance_matrix_masked(filaments: np.ndarray, mask: PairMask, with_methods: bool = False):
    """partial_inductance_matrix_masked (inductance/batch.rs:305-312): pairs
    the mask rejects are exactly 0 and never evaluated (method -1)."""
    f = np.ascontiguousarray(filaments, dtype=FILAMENT_DTYPE)
    n = len(f)
    group = np.ascontiguousarray(mask.group, dtype=np.int32)
    table = np.ascontiguousarray(mask.coupled, dtype=np.uint8)
    pm = native.PairMask(native.ptr(group) if len(group) else None, len(group), mask.groups,
                         native.ptr(table) if table.size else None)
    L = np.zeros((n, n))
    methods = np.zeros((n, n), dtype=np.int8) if with_methods else None
    native.check(native.lib().fh_partial_inductance_masked(
        native.ptr(f), n, C.byref(pm), native.ptr(L), native.ptr(methods) if with_methods else None))
    return (L, methods) if with_methods else L


def partial_inductance_matrix(filaments: np.ndarray, with_methods: bool = False):
    """The symmetric partial-inductance matrix on the GPU (batch.rs:246-253)."""
    f = np.ascontiguousarray(filaments, dtype=FILAMENT_DTYPE)
    n = len(f)
    L = np.zeros((n, n))
    methods = np.zeros((n, n), dtype=np.int8) if with_methods else None
    native.check(native.lib().fh_partial_inductance(
        native.ptr(f), n, native.ptr(L), native.ptr(methods) if with_methods else None))
    return (L, methods) if with_methods else L


def device_math(x) -> tuple:
    """log and atan as the fill kernels evaluate them on the device."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    lg = np.zeros_like(x)
    at = np.zeros_like(x)
    native.check(native.lib().fh_device_math(native.ptr(x), len(x), native.ptr(lg), native.ptr(at)))
    return lg, at


def version() -> str:
    return native.lib().fh_version().decode()
