"""ctypes binding of libfasterhenry_gpu.so (include/fasterhenry_gpu.h).

The library is built in-tree (fasterhenry_amd/lib) by
``__graft_entry__.build()`` or ``make -C fasterhenry_amd/csrc``.  There is no
fallback: if the library is missing, or no HIP device is visible, every
entry raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import errors

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libfasterhenry_gpu.so")

FH_OK = 0
FH_FP64_PLAIN = 0
FH_FP32_REFINED = 1
FH_DENSE_PATH_MAX_FILAMENTS = 10000
FH_REFINE_SPLIT_CLASSICAL = 1
FH_REFINE_SPLIT_GMRES = 2
FH_REFINE_FP32_CLASSICAL = 4
FH_REFINE_FP32_GMRES = 8
FH_REFINE_FP64 = 16
FH_REFINE_ALL = 31

SYMBOLS = [
    "fh_resolve", "fh_system_create", "fh_sweep", "fh_impedance", "fh_system_set_precision",
    "fh_system_set_profiling", "fh_system_counts", "fh_system_filaments",
    "fh_system_inductance", "fh_system_blocks", "fh_system_destroy", "fh_solve",
    "fh_partial_inductance", "fh_discretize", "fh_device_math", "fh_last_error",
    "fh_error_detail", "fh_version", "fh_iterative_default_params", "fh_system_create_iterative",
    "fh_system_operator_stats", "fh_pfft_default_params", "fh_pfft_create", "fh_pfft_apply",
    "fh_pfft_get_stats", "fh_pfft_data", "fh_pfft_destroy", "fh_gmres_dense", "fh_solve_batch",
    "fh_discretize_tangent", "fh_grad", "fh_partial_inductance_derivative", "fh_currents",
    "fh_zc_mat", "fh_spice_subckt", "fh_subckt_name_from_path", "fh_system_create_coupled",
    "fh_partial_inductance_masked", "fh_deck_parse", "fh_deck_problem", "fh_deck_info",
    "fh_deck_coupling", "fh_deck_title", "fh_deck_port_name", "fh_deck_group_name", "fh_deck_destroy",
    "fh_system_set_sweep_parts", "fh_partial_inductance_noalloc", "fh_system_set_refinement",
    "fh_system_get_schedule", "fh_system_set_graphs", "fh_system_set_schedule", "fh_system_autotune", "fh_tune_save",
    "fh_tune_load",
]

_ERRORS = {
    1: errors.InvalidArgument, 2: errors.GeometryError, 3: errors.DiscretizeError,
    4: errors.KernelError, 5: errors.MeshError, 6: errors.InvalidGrading,
    7: errors.InvalidFrequency, 8: errors.Singular, 9: errors.NotConverged,
    10: errors.SolveError, 11: errors.SolveError, 12: errors.HipError, 13: errors.PfftError,
    14: errors.InvalidExcitation, 15: errors.DeckError,
}


class Problem(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("n_nodes", C.c_size_t), ("segments", C.c_void_p),
                ("grids", C.c_void_p), ("n_segments", C.c_size_t), ("ports", C.c_void_p),
                ("n_ports", C.c_size_t)]


class SweepInfo(C.Structure):
    _fields_ = [("nodes", C.c_size_t), ("segments", C.c_size_t), ("filaments", C.c_size_t),
                ("meshes", C.c_size_t), ("internal_meshes", C.c_size_t), ("ports", C.c_size_t),
                ("inductance_s", C.c_double), ("assembly_s", C.c_double),
                ("solve_s", C.c_double), ("refine_steps_max", C.c_int32),
                ("factor_arith", C.c_int32), ("refine_change_max", C.c_double),
                ("update_kernel_ms", C.c_double), ("update_kernel_launches", C.c_int64),
                ("update_kernel_flops", C.c_double), ("update_kernel_unit_flops", C.c_double),
                ("solver", C.c_int32), ("gmres_iterations_max", C.c_int32),
                ("gmres_iterations", C.c_int64), ("gmres_residual_max", C.c_double),
                ("operator_applies", C.c_int64), ("operator_apply_ms", C.c_double),
                ("refine_stage", C.c_int32)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Schedule(C.Structure):
    """fh_schedule: outer block of the factorization and sweep parts."""
    _fields_ = [("outer_block", C.c_int32), ("sweep_parts", C.c_int32)]

    def as_dict(self) -> dict:
        return {"outer_block": self.outer_block, "sweep_parts": self.sweep_parts}


class PairMask(C.Structure):
    """fh_pair_mask: PairMask (inductance/batch.rs:37-110)."""
    _fields_ = [("group", C.c_void_p), ("n", C.c_size_t), ("groups", C.c_size_t),
                ("coupled", C.c_void_p)]


class CouplingMask(C.Structure):
    """fh_coupling: a Coupling resolved against the segments."""
    _fields_ = [("segment_group", C.c_void_p), ("n_segments", C.c_size_t), ("groups", C.c_size_t),
                ("coupled", C.c_void_p)]


class PowerBalance(C.Structure):
    """fh_power_balance: PowerBalance (currents.rs:103-111)."""
    _fields_ = [("dissipated_w", C.c_double), ("input_w", C.c_double),
                ("relative_residual", C.c_double)]


class ParamMap(C.Structure):
    """fh_param_map: geometry directions of the gradient parameters."""
    _fields_ = [("n_params", C.c_size_t), ("d_nodes", C.c_void_p), ("d_widths", C.c_void_p),
                ("d_heights", C.c_void_p)]


class PfftParams(C.Structure):
    """fh_pfft_params: PfftParams (pfft/mod.rs:183-231)."""
    _fields_ = [("auto_grid", C.c_int32), ("interpolation_order", C.c_int32),
                ("cells_per_filament", C.c_double), ("spacing", C.c_double),
                ("near_field_radius", C.c_double), ("fft_precision", C.c_int32),
                ("reserved", C.c_int32)]


class GmresParams(C.Structure):
    """fh_gmres_params: GmresParams (gmres.rs:29-60)."""
    _fields_ = [("tolerance", C.c_double), ("restart", C.c_uint64),
                ("max_iterations", C.c_uint64)]


class IterativeParams(C.Structure):
    """fh_iterative_params: IterativeParams (iterative.rs:98-105)."""
    _fields_ = [("pfft", PfftParams), ("gmres", GmresParams)]


class PfftStats(C.Structure):
    """fh_pfft_stats: PfftStats (pfft/mod.rs:255-270)."""
    _fields_ = [("filaments", C.c_uint64), ("grid_spacing", C.c_double),
                ("grid_dims", C.c_uint64 * 3), ("fft_dims", C.c_uint64 * 3),
                ("projection_entries", C.c_uint64), ("near_entries", C.c_uint64),
                ("memory_bytes", C.c_uint64), ("fft_kind", C.c_int32),
                ("line_group", C.c_int32 * 3)]

    def as_dict(self) -> dict:
        return {"filaments": self.filaments, "grid_spacing": self.grid_spacing,
                "grid_dims": list(self.grid_dims), "fft_dims": list(self.fft_dims),
                "projection_entries": self.projection_entries,
                "near_entries": self.near_entries, "memory_bytes": self.memory_bytes,
                "fft_kind": self.fft_kind, "line_group": list(self.line_group)}


_lib = None


def lib() -> C.CDLL:
    """The loaded native library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() or "
                "`make -C fasterhenry_amd/csrc`; there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        vp, sz, dp = C.c_void_p, C.c_size_t, C.POINTER(C.c_double)
        st = C.c_int
        L.fh_resolve.restype = C.c_int
        L.fh_resolve.argtypes = [C.c_int, sz]
        L.fh_system_create.restype = st
        L.fh_system_create.argtypes = [C.POINTER(Problem), C.c_int, C.c_int, C.POINTER(vp)]
        L.fh_sweep.restype = st
        L.fh_sweep.argtypes = [vp, vp, sz, vp, C.POINTER(SweepInfo)]
        L.fh_impedance.restype = st
        L.fh_impedance.argtypes = [vp, C.c_double, vp]
        L.fh_system_set_precision.restype = st
        L.fh_system_set_precision.argtypes = [vp, C.c_int]
        L.fh_system_set_profiling.restype = st
        L.fh_system_set_profiling.argtypes = [vp, C.c_int]
        L.fh_system_counts.restype = st
        L.fh_system_counts.argtypes = [vp, C.POINTER(sz)]
        L.fh_system_filaments.restype = st
        L.fh_system_filaments.argtypes = [vp, vp]
        L.fh_system_inductance.restype = st
        L.fh_system_inductance.argtypes = [vp, vp]
        L.fh_system_blocks.restype = st
        L.fh_system_blocks.argtypes = [vp] * 7
        L.fh_system_destroy.restype = None
        L.fh_system_destroy.argtypes = [vp]
        L.fh_solve.restype = st
        L.fh_solve.argtypes = [C.POINTER(Problem), C.c_int, C.c_int, vp, sz, vp,
                               C.POINTER(SweepInfo)]
        L.fh_partial_inductance.restype = st
        L.fh_partial_inductance.argtypes = [vp, sz, vp, vp]
        L.fh_system_set_graphs.restype = st
        L.fh_system_set_graphs.argtypes = [vp, C.c_int32]
        L.fh_system_get_schedule.restype = st
        L.fh_system_get_schedule.argtypes = [vp, C.POINTER(Schedule)]
        L.fh_system_set_schedule.restype = st
        L.fh_system_set_schedule.argtypes = [vp, C.POINTER(Schedule)]
        L.fh_system_autotune.restype = st
        L.fh_system_autotune.argtypes = [vp, vp, sz, C.POINTER(Schedule), C.POINTER(C.c_double)]
        L.fh_tune_save.restype = st
        L.fh_tune_save.argtypes = [C.c_char_p]
        L.fh_tune_load.restype = st
        L.fh_tune_load.argtypes = [C.c_char_p]
        L.fh_system_set_refinement.restype = st
        L.fh_system_set_refinement.argtypes = [vp, C.c_uint32]
        L.fh_partial_inductance_noalloc.restype = st
        L.fh_partial_inductance_noalloc.argtypes = [vp, sz, vp, vp, vp]
        L.fh_discretize.restype = st
        L.fh_discretize.argtypes = [C.POINTER(Problem), vp, sz, C.POINTER(sz)]
        L.fh_device_math.restype = st
        L.fh_device_math.argtypes = [vp, sz, vp, vp]
        L.fh_last_error.restype = C.c_char_p
        L.fh_last_error.argtypes = []
        L.fh_error_detail.restype = None
        L.fh_error_detail.argtypes = [C.POINTER(C.c_int64), C.POINTER(C.c_int64), dp,
                                      C.POINTER(C.c_int32)]
        L.fh_version.restype = C.c_char_p
        L.fh_iterative_default_params.restype = None
        L.fh_iterative_default_params.argtypes = [C.POINTER(IterativeParams)]
        L.fh_system_create_iterative.restype = st
        L.fh_system_create_iterative.argtypes = [C.POINTER(Problem), C.POINTER(IterativeParams),
                                                 C.POINTER(vp)]
        L.fh_system_operator_stats.restype = st
        L.fh_system_operator_stats.argtypes = [vp, C.POINTER(PfftStats)]
        L.fh_pfft_default_params.restype = None
        L.fh_pfft_default_params.argtypes = [C.POINTER(PfftParams)]
        L.fh_pfft_create.restype = st
        L.fh_pfft_create.argtypes = [vp, sz, C.POINTER(PfftParams), C.POINTER(vp)]
        L.fh_pfft_apply.restype = st
        L.fh_pfft_apply.argtypes = [vp, vp, vp, sz]
        L.fh_pfft_get_stats.restype = st
        L.fh_pfft_get_stats.argtypes = [vp, C.POINTER(PfftStats)]
        L.fh_pfft_data.restype = st
        L.fh_pfft_data.argtypes = [vp] * 7
        L.fh_pfft_destroy.restype = None
        L.fh_pfft_destroy.argtypes = [vp]
        L.fh_discretize_tangent.restype = st
        L.fh_discretize_tangent.argtypes = [C.POINTER(Problem), vp, vp, vp, vp, sz]
        L.fh_grad.restype = st
        L.fh_grad.argtypes = [vp, C.POINTER(ParamMap), vp, sz, vp, vp, vp]
        L.fh_system_create_coupled.restype = st
        L.fh_system_create_coupled.argtypes = [C.POINTER(Problem), C.c_int, C.c_int,
                                               C.POINTER(CouplingMask), C.POINTER(vp)]
        L.fh_partial_inductance_masked.restype = st
        L.fh_partial_inductance_masked.argtypes = [vp, sz, C.POINTER(PairMask), vp, vp]
        L.fh_system_set_sweep_parts.restype = st
        L.fh_system_set_sweep_parts.argtypes = [vp, C.c_int32]
        L.fh_deck_parse.restype = st
        L.fh_deck_parse.argtypes = [C.c_char_p, sz, C.POINTER(vp)]
        L.fh_deck_problem.restype = st
        L.fh_deck_problem.argtypes = [vp, C.POINTER(Problem)]
        L.fh_deck_info.restype = st
        L.fh_deck_info.argtypes = [vp, C.POINTER(sz), C.POINTER(C.POINTER(C.c_double)),
                                   C.POINTER(C.c_int32), C.POINTER(sz)]
        L.fh_deck_coupling.restype = st
        L.fh_deck_coupling.argtypes = [vp, C.POINTER(CouplingMask)]
        for name in ("fh_deck_title", "fh_deck_port_name", "fh_deck_group_name"):
            getattr(L, name).restype = C.c_char_p
        L.fh_deck_title.argtypes = [vp]
        L.fh_deck_port_name.argtypes = [vp, sz]
        L.fh_deck_group_name.argtypes = [vp, sz]
        L.fh_deck_destroy.restype = None
        L.fh_deck_destroy.argtypes = [vp]
        L.fh_zc_mat.restype = st
        L.fh_zc_mat.argtypes = [vp, sz, vp, sz, vp, sz, C.POINTER(sz)]
        L.fh_spice_subckt.restype = st
        L.fh_spice_subckt.argtypes = [vp, sz, vp, sz, sz, C.c_char_p, C.c_char_p, sz, C.POINTER(sz)]
        L.fh_subckt_name_from_path.restype = st
        L.fh_subckt_name_from_path.argtypes = [C.c_char_p, C.c_char_p, sz, C.POINTER(sz)]
        L.fh_currents.restype = st
        L.fh_currents.argtypes = [vp, C.c_double, vp, sz, vp, vp, vp, vp, vp, vp, C.POINTER(sz), vp, vp,
                                  C.POINTER(PowerBalance)]
        L.fh_partial_inductance_derivative.restype = st
        L.fh_partial_inductance_derivative.argtypes = [vp, sz, vp, sz, vp, sz, vp, vp]
        L.fh_solve_batch.restype = st
        L.fh_solve_batch.argtypes = [vp, sz, C.c_int, C.c_int, vp, sz, vp, vp, C.c_int32,
                                     C.POINTER(sz)]
        L.fh_gmres_dense.restype = st
        L.fh_gmres_dense.argtypes = [sz, vp, vp, vp, C.POINTER(GmresParams), C.c_int32,
                                     C.POINTER(C.c_uint64), dp, C.POINTER(C.c_int32)]
        _lib = L
    return _lib


def check(status: int) -> None:
    """Raise the SolveError subclass matching a non-zero fh_status."""
    if status == FH_OK:
        return
    L = lib()
    msg = L.fh_last_error().decode(errors="replace")
    i0, i1, v, code = C.c_int64(), C.c_int64(), C.c_double(), C.c_int32()
    L.fh_error_detail(C.byref(i0), C.byref(i1), C.byref(v), C.byref(code))
    cls = _ERRORS.get(status, errors.SolveError)
    raise cls(msg or f"fh_status {status}", status=status, index0=i0.value, index1=i1.value,
              value=v.value, code=code.value)


def ptr(a: np.ndarray) -> C.c_void_p:
    return a.ctypes.data_as(C.c_void_p)


def problem_struct(packed) -> Problem:
    p = Problem()
    p.nodes = ptr(packed.nodes) if len(packed.nodes) else None
    p.n_nodes = len(packed.nodes)
    p.segments = ptr(packed.segments) if len(packed.segments) else None
    p.grids = ptr(packed.grids) if len(packed.grids) else None
    p.n_segments = len(packed.segments)
    p.ports = ptr(packed.ports) if len(packed.ports) else None
    p.n_ports = len(packed.ports)
    return p
