"""ctypes binding of the oracle (oracle/fh_oracle.c).

TEST INFRASTRUCTURE ONLY: the CPU FP64 restatement of the reference's hot
path that tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
use as the checker.  The product (fasterhenry_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libfh_oracle.so")

METHODS = {0: "orthogonal", 1: "bar_closed_form", 2: "aligned_quadrature",
           3: "point_quadrature", 4: "sampled_filaments"}

FHO_OK = 0
FHO_E_KERNEL = 4
FHO_E_MESH = 5
FHO_E_INVALID_FREQUENCY = 7
FHO_E_SINGULAR = 8


class OracleError(RuntimeError):
    def __init__(self, status: int, detail: "Detail"):
        super().__init__(f"oracle status {status}: code {detail.code} "
                         f"index ({detail.index0}, {detail.index1}) value {detail.value}")
        self.status = status
        self.detail = detail


class Detail(C.Structure):
    _fields_ = [("index0", C.c_int64), ("index1", C.c_int64), ("value", C.c_double),
                ("code", C.c_int)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"{_LIB_PATH} is missing: run `make -C oracle` "
                               "(or __graft_entry__.build())")
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        vp = C.c_void_p
        sz = C.c_size_t
        L.fho_set_math.argtypes = [C.c_int]
        L.fho_log.restype = C.c_double
        L.fho_log.argtypes = [C.c_double]
        L.fho_atan.restype = C.c_double
        L.fho_atan.argtypes = [C.c_double]
        L.fho_hypot.restype = C.c_double
        L.fho_hypot.argtypes = [C.c_double, C.c_double]
        L.fho_last_error.argtypes = [C.POINTER(Detail)]
        L.fho_gauss_rule.argtypes = [C.c_int, dp, dp]
        L.fho_order_for.restype = C.c_int
        L.fho_order_for.argtypes = [C.c_double, C.c_int, C.c_double, C.c_double]
        L.fho_bar_integral.restype = C.c_double
        L.fho_bar_integral.argtypes = [dp, dp, dp, dp]
        L.fho_corner_differences.argtypes = [C.c_double, C.c_double, C.c_double, dp]
        L.fho_lines_parallel.restype = C.c_int
        L.fho_lines_parallel.argtypes = [C.c_double] * 4 + [dp]
        L.fho_lines_skew.restype = C.c_int
        L.fho_lines_skew.argtypes = [dp, dp, dp, dp, C.c_double, dp]
        L.fho_aligned_integral.restype = C.c_double
        L.fho_aligned_integral.argtypes = [dp]
        L.fho_aligned_dispatch.restype = C.c_double
        L.fho_aligned_dispatch.argtypes = [dp, C.POINTER(C.c_int)]
        L.fho_discretize_geometry.restype = C.c_int
        L.fho_discretize_geometry.argtypes = [vp, sz, vp, sz, vp, vp, sz, C.POINTER(sz)]
        L.fho_mutual.restype = C.c_int
        L.fho_mutual.argtypes = [vp, vp, dp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.fho_self_inductance.restype = C.c_double
        L.fho_self_inductance.argtypes = [vp]
        L.fho_fill.restype = C.c_int
        L.fho_fill.argtypes = [vp, sz, vp, C.c_int]
        L.fho_fill_pairs.restype = C.c_int
        L.fho_fill_pairs.argtypes = [vp, sz, vp, vp, sz, vp, vp, C.c_int]
        L.fho_mesh_build.restype = C.c_int
        L.fho_mesh_build.argtypes = [sz, vp, sz, vp, vp, sz, vp, vp, vp, C.POINTER(sz),
                                     C.POINTER(sz), C.POINTER(sz)]
        L.fho_congruence.restype = C.c_int
        L.fho_congruence.argtypes = [sz, vp, vp, vp, sz, vp, vp, vp, C.c_int]
        L.fho_lu_solve_complex.restype = C.c_int
        L.fho_lu_solve_complex.argtypes = [vp, sz, vp, sz, C.c_int]
        L.fho_lu_solve_real.restype = C.c_int
        L.fho_lu_solve_real.argtypes = [vp, sz, vp, sz, C.c_int]
        L.fho_system_create.restype = C.c_int
        L.fho_system_create.argtypes = [vp, sz, vp, sz, vp, vp, sz, C.c_int, C.POINTER(vp)]
        L.fho_system_destroy.argtypes = [vp]
        L.fho_system_counts.argtypes = [vp, C.POINTER(sz)]
        L.fho_system_blocks.restype = C.c_int
        L.fho_system_blocks.argtypes = [vp] * 7
        L.fho_system_inductance.restype = vp
        L.fho_system_inductance.argtypes = [vp]
        L.fho_system_filaments.restype = vp
        L.fho_system_filaments.argtypes = [vp]
        L.fho_system_sweep.restype = C.c_int
        L.fho_system_sweep.argtypes = [vp, vp, sz, vp, C.c_int]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _check(status: int) -> None:
    if status != FHO_OK:
        d = Detail()
        lib().fho_last_error(C.byref(d))
        raise OracleError(status, d)


def set_portable_math(portable: bool) -> None:
    lib().fho_set_math(1 if portable else 0)


def log(x: float) -> float:
    return lib().fho_log(x)


def atan(x: float) -> float:
    return lib().fho_atan(x)


def gauss_rule(n: int) -> Tuple[np.ndarray, np.ndarray]:
    nodes = np.zeros(max(n, 1))
    weights = np.zeros(max(n, 1))
    dp = C.POINTER(C.c_double)
    lib().fho_gauss_rule(n, nodes.ctypes.data_as(dp), weights.ctypes.data_as(dp))
    return nodes, weights


def order_for(tolerance: float, max_order: int, half_extent: float, distance: float):
    r = lib().fho_order_for(tolerance, max_order, half_extent, distance)
    return None if r == 0 else r


def corner_differences(e1: float, e2: float, offset: float) -> np.ndarray:
    out = np.zeros(4)
    lib().fho_corner_differences(e1, e2, offset, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def bar_integral(qx, qy, qz) -> Tuple[float, float]:
    dp = C.POINTER(C.c_double)
    a, b, c = (np.ascontiguousarray(q, dtype=np.float64) for q in (qx, qy, qz))
    err = C.c_double()
    v = lib().fho_bar_integral(a.ctypes.data_as(dp), b.ctypes.data_as(dp), c.ctypes.data_as(dp),
                               C.byref(err))
    return v, err.value


def lines_parallel(l1: float, l2: float, offset: float, rho: float):
    out = C.c_double()
    ok = lib().fho_lines_parallel(l1, l2, offset, rho, C.byref(out))
    return out.value if ok else None


def lines_skew(a0, a1, b0, b1, sin_min: float = 0.0):
    dp = C.POINTER(C.c_double)
    arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (a0, a1, b0, b1)]
    out = C.c_double()
    ok = lib().fho_lines_skew(*(x.ctypes.data_as(dp) for x in arrs), sin_min, C.byref(out))
    return out.value if ok else None


def aligned_integral(bars) -> float:
    b = np.ascontiguousarray(bars, dtype=np.float64)
    return lib().fho_aligned_integral(b.ctypes.data_as(C.POINTER(C.c_double)))


def aligned_dispatch(bars) -> Tuple[float, str]:
    b = np.ascontiguousarray(bars, dtype=np.float64)
    m = C.c_int()
    v = lib().fho_aligned_dispatch(b.ctypes.data_as(C.POINTER(C.c_double)), C.byref(m))
    return v, METHODS[m.value]


def discretize(packed) -> np.ndarray:
    """Filaments (FILAMENT_DTYPE) of a fasterhenry_amd.model.PackedProblem."""
    from fasterhenry_amd.model import FILAMENT_DTYPE

    n = C.c_size_t()
    L = lib()
    _check(L.fho_discretize_geometry(_ptr(packed.nodes), len(packed.nodes), _ptr(packed.segments),
                                     len(packed.segments), _ptr(packed.grids), None, 0,
                                     C.byref(n)))
    out = np.zeros(n.value, dtype=FILAMENT_DTYPE)
    _check(L.fho_discretize_geometry(_ptr(packed.nodes), len(packed.nodes), _ptr(packed.segments),
                                     len(packed.segments), _ptr(packed.grids), _ptr(out), n.value,
                                     C.byref(n)))
    return out


def mutual(fa: np.ndarray, fb: np.ndarray) -> Tuple[float, str, bool]:
    """mutual_inductance_detailed of two filaments (FILAMENT_DTYPE scalars)."""
    a = np.array(fa).reshape(1)
    b = np.array(fb).reshape(1)
    v = C.c_double()
    m = C.c_int()
    r = C.c_int()
    _check(lib().fho_mutual(_ptr(a), _ptr(b), C.byref(v), C.byref(m), C.byref(r)))
    return v.value, METHODS[m.value], bool(r.value)


def self_inductance(f: np.ndarray) -> float:
    a = np.array(f).reshape(1)
    return lib().fho_self_inductance(_ptr(a))


def fill(filaments: np.ndarray, threads: int = 0) -> np.ndarray:
    n = len(filaments)
    L = np.zeros((n, n))
    _check(lib().fho_fill(_ptr(filaments), n, _ptr(L), threads))
    return L


def fill_pairs(filaments: np.ndarray, rows, cols, threads: int = 0):
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    out = np.zeros(len(rows))
    methods = np.zeros(len(rows), dtype=np.int32)
    _check(lib().fho_fill_pairs(_ptr(filaments), len(filaments), _ptr(rows), _ptr(cols),
                                len(rows), _ptr(out), _ptr(methods), threads))
    return out, methods


def mesh_build(packed):
    """CSR loop basis: (row_ptr, branch, sign, internal)."""
    L = lib()
    per = (packed.grids["nw"].astype(np.uint64) * packed.grids["nh"]).astype(np.uint64)
    per = np.ascontiguousarray(per, dtype=np.uintp)
    loops, internal, nnz = C.c_size_t(), C.c_size_t(), C.c_size_t()
    _check(L.fho_mesh_build(len(packed.nodes), _ptr(packed.segments), len(packed.segments),
                            _ptr(per), _ptr(packed.ports), len(packed.ports), None, None, None,
                            C.byref(loops), C.byref(internal), C.byref(nnz)))
    row_ptr = np.zeros(loops.value + 1, dtype=np.int64)
    branch = np.zeros(max(nnz.value, 1), dtype=np.int64)
    sign = np.zeros(max(nnz.value, 1))
    _check(L.fho_mesh_build(len(packed.nodes), _ptr(packed.segments), len(packed.segments),
                            _ptr(per), _ptr(packed.ports), len(packed.ports), _ptr(row_ptr),
                            _ptr(branch), _ptr(sign), C.byref(loops), C.byref(internal),
                            C.byref(nnz)))
    return row_ptr, branch[:nnz.value], sign[:nnz.value], internal.value


def lu_solve(a: np.ndarray, b: np.ndarray, threads: int = 0) -> np.ndarray:
    """dense.rs lu_solve: x = a^-1 b (complex or real)."""
    complex_ = np.iscomplexobj(a) or np.iscomplexobj(b)
    dt = np.complex128 if complex_ else np.float64
    A = np.asfortranarray(a, dtype=dt).copy(order="F")
    B = np.asfortranarray(b.reshape(len(b), -1), dtype=dt).copy(order="F")
    f = lib().fho_lu_solve_complex if complex_ else lib().fho_lu_solve_real
    _check(f(_ptr(A), A.shape[0], _ptr(B), B.shape[1], threads))
    return B.reshape(b.shape)


class System:
    """An assembled oracle mesh system (solve.rs MeshSystem)."""

    def __init__(self, packed, threads: int = 0):
        self._h = C.c_void_p()
        st = lib().fho_system_create(_ptr(packed.nodes), len(packed.nodes), _ptr(packed.segments),
                                     len(packed.segments), _ptr(packed.grids), _ptr(packed.ports),
                                     len(packed.ports), threads, C.byref(self._h))
        _check(st)
        self.threads = threads
        c = (C.c_size_t * 6)()
        lib().fho_system_counts(self._h, c)
        (self.nodes, self.segments, self.filaments, self.meshes, self.internal,
         self.ports) = list(c)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().fho_system_destroy(self._h)
            self._h = None

    @property
    def counts(self) -> dict:
        return dict(nodes=self.nodes, segments=self.segments, filaments=self.filaments,
                    meshes=self.meshes, internal_meshes=self.internal, ports=self.ports)

    def blocks(self):
        m, p = self.internal, self.ports
        out = [np.zeros((m, m)), np.zeros((m, m)), np.zeros((m, p)), np.zeros((m, p)),
               np.zeros((p, p)), np.zeros((p, p))]
        lib().fho_system_blocks(self._h, *(_ptr(x) for x in out))
        r_ee, l_ee, r_ep, l_ep, r_pp, l_pp = out
        return dict(r_ee=r_ee, l_ee=l_ee, r_ep=r_ep, l_ep=l_ep, r_pp=r_pp, l_pp=l_pp)

    def inductance(self) -> np.ndarray:
        n = self.filaments
        ptr = lib().fho_system_inductance(self._h)
        return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_double)), shape=(n, n)).copy()

    def filament_array(self) -> np.ndarray:
        from fasterhenry_amd.model import FILAMENT_DTYPE

        ptr = lib().fho_system_filaments(self._h)
        raw = (C.c_char * (self.filaments * FILAMENT_DTYPE.itemsize)).from_address(ptr)
        return np.frombuffer(bytes(raw), dtype=FILAMENT_DTYPE).copy()

    def sweep(self, freqs) -> np.ndarray:
        f = np.ascontiguousarray(freqs, dtype=np.float64)
        z = np.zeros((len(f), self.ports, self.ports), dtype=np.complex128)
        _check(lib().fho_system_sweep(self._h, _ptr(f), len(f), _ptr(z), self.threads))
        return z


def solve(packed, freqs, threads: int = 0) -> np.ndarray:
    return System(packed, threads).sweep(freqs)


def currents(packed, frequency: float, excitation, system: "System" = None) -> dict:
    """solve_currents (solve.rs:1110-1149) and assemble_solution
    (currents.rs:190-274) on the oracle's mesh blocks: internal loop
    currents -Zee^-1 (Zep I_p) by the oracle's LU, port voltages
    Zpp I_p + Zep^T I_e, branch currents M^T [I_e; I_p], then |J|, power,
    segment sums, conductors (Layout::new, currents.rs:144-184) and the
    power balance.  Returns a dict of numpy arrays and scalars."""
    system = system or System(packed)
    ex = np.asarray(excitation, dtype=np.complex128).reshape(-1)
    b = system.blocks()
    m, p = system.internal, system.ports
    w = 2.0 * np.pi * frequency
    ep = b["r_ep"] + 1j * w * b["l_ep"]
    pp = b["r_pp"] + 1j * w * b["l_pp"]
    if m:
        ee = b["r_ee"] + 1j * w * b["l_ee"]
        internal = -lu_solve(ee, ep @ ex)
    else:
        internal = np.zeros(0, dtype=np.complex128)
    voltage = pp @ ex + ep.T @ internal
    row_ptr, branch, sign, _ = mesh_build(packed)
    f = system.filament_array()
    n = len(f)
    loops = np.concatenate([internal, ex])
    cur = np.zeros(n, dtype=np.complex128)
    for l in range(len(loops)):
        for e in range(row_ptr[l], row_ptr[l + 1]):
            cur[branch[e]] += loops[l] * sign[e]
    width, height = np.asarray(f["width"]), np.asarray(f["height"])
    res = np.asarray(f["length"]) / (np.asarray(f["sigma"]) * (width * height))
    density = np.abs(cur) / (width * height)
    power = 0.5 * np.abs(cur) ** 2 * res
    nodes = len(packed.nodes)
    parent = list(range(nodes))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    segs = packed.segments
    for q in range(len(segs)):
        a, c = find(int(segs[q]["a"])), find(int(segs[q]["b"]))
        if a != c:
            parent[a] = c
    roots, conductor = [], []
    for q in range(len(segs)):
        r = find(int(segs[q]["a"]))
        if r not in roots:
            roots.append(r)
        conductor.append(roots.index(r))
    per = packed.grids["nw"].astype(np.int64) * packed.grids["nh"].astype(np.int64)
    seg_of = np.repeat(np.arange(len(segs)), per)
    seg_current = np.zeros(len(segs), dtype=np.complex128)
    seg_power = np.zeros(len(segs))
    peak = {}
    dissipated = 0.0
    for k in range(n):
        q = seg_of[k]
        dissipated += power[k]
        seg_current[q] += cur[k]
        seg_power[q] += power[k]
        c = conductor[q]
        if c not in peak or density[k] > density[peak[c]]:
            peak[c] = k
    inp = 0.5 * float(np.sum((np.conj(ex) * voltage).real))
    scale = max(dissipated, inp)
    return dict(current=cur, density=density, power=power, segment_current=seg_current,
                segment_power=seg_power, conductor=np.array(conductor),
                peaks=[peak[c] for c in range(len(roots))], voltage=voltage,
                location=(np.asarray(f["start"]) + np.asarray(f["end"])) * 0.5,
                dissipated_w=dissipated, input_w=inp,
                relative_residual=abs(dissipated - inp) / scale if scale > 0 else 0.0)


# ---------------------------------------------------------------------------
# pfft/*.rs, gmres.rs, iterative.rs

FHO_E_NOT_CONVERGED = 9


class PfftParams(C.Structure):
    """PfftParams (pfft/mod.rs:183-231); defaults p = 5, k = 8, 8 cells per
    filament."""
    _fields_ = [("auto_grid", C.c_int), ("cells_per_filament", C.c_double),
                ("spacing", C.c_double), ("near_field_radius", C.c_double),
                ("interpolation_order", C.c_int)]

    @classmethod
    def default(cls, spacing: float = None, near_field_radius: float = 8.0, order: int = 5,
                cells_per_filament: float = 8.0) -> "PfftParams":
        p = cls()
        p.auto_grid = 0 if spacing is not None else 1
        p.cells_per_filament = cells_per_filament
        p.spacing = spacing or 0.0
        p.near_field_radius = near_field_radius
        p.interpolation_order = order
        return p


_pfft_bound = False


def _bind_pfft() -> C.CDLL:
    global _pfft_bound
    L = lib()
    if not _pfft_bound:
        vp, sz, dp = C.c_void_p, C.c_size_t, C.POINTER(C.c_double)
        L.fho_pfft_create.restype = C.c_int
        L.fho_pfft_create.argtypes = [vp, sz, C.POINTER(PfftParams), C.POINTER(vp)]
        L.fho_pfft_destroy.argtypes = [vp]
        L.fho_pfft_apply.argtypes = [vp, vp, vp]
        L.fho_pfft_stats.argtypes = [vp, dp, vp, vp, C.POINTER(sz), C.POINTER(sz)]
        L.fho_pfft_data.restype = C.c_int
        L.fho_pfft_data.argtypes = [vp] * 8
        L.fho_pfft_convolve.restype = C.c_int
        L.fho_pfft_convolve.argtypes = [vp] * 5
        L.fho_auto_spacing.restype = C.c_double
        L.fho_auto_spacing.argtypes = [vp, C.c_double]
        L.fho_smooth_size.restype = sz
        L.fho_smooth_size.argtypes = [sz]
        L.fho_axis_stencil.restype = sz
        L.fho_axis_stencil.argtypes = [vp, C.c_double, vp, C.c_int, C.c_int, C.c_double, vp]
        L.fho_gmres_dense.restype = C.c_int
        L.fho_gmres_dense.argtypes = [sz, vp, vp, vp, C.c_double, sz, sz, C.c_int,
                                      C.POINTER(sz), dp, C.POINTER(C.c_int)]
        L.fho_iter_create.restype = C.c_int
        L.fho_iter_create.argtypes = [vp, sz, vp, sz, vp, vp, sz, C.POINTER(PfftParams),
                                      C.c_double, sz, sz, C.POINTER(vp)]
        L.fho_iter_destroy.argtypes = [vp]
        L.fho_iter_impedance.restype = C.c_int
        L.fho_iter_impedance.argtypes = [vp, C.c_double, vp, vp]
        L.fho_iter_counts.argtypes = [vp, C.POINTER(sz)]
        L.fho_iter_operator.restype = vp
        L.fho_iter_operator.argtypes = [vp]
        _pfft_bound = True
    return L


def auto_spacing(extent, cells: float) -> float:
    e = np.ascontiguousarray(extent, dtype=np.float64)
    return _bind_pfft().fho_auto_spacing(_ptr(e), cells)


def smooth_size(n: int) -> int:
    return _bind_pfft().fho_smooth_size(n)


def axis_stencil(origin, h: float, dims, m: int, axis: int, x: float):
    o = np.ascontiguousarray(origin, dtype=np.float64)
    d = np.ascontiguousarray(dims, dtype=np.uintp)
    w = np.zeros(m)
    first = _bind_pfft().fho_axis_stencil(_ptr(o), h, _ptr(d), m, axis, x, _ptr(w))
    return first, w


class Pfft:
    """PfftOperator (pfft/mod.rs:270-528) of a FILAMENT_DTYPE array."""

    def __init__(self, filaments: np.ndarray, params: PfftParams = None):
        L = _bind_pfft()
        self._f = np.ascontiguousarray(filaments)
        self.n = len(self._f)
        self.params = params or PfftParams.default()
        self._h = C.c_void_p()
        _check(L.fho_pfft_create(_ptr(self._f), self.n, C.byref(self.params), C.byref(self._h)))

    def __del__(self):
        if getattr(self, "_h", None):
            _bind_pfft().fho_pfft_destroy(self._h)
            self._h = None

    def apply(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n)
        _bind_pfft().fho_pfft_apply(self._h, _ptr(x), _ptr(y))
        return y

    def stats(self) -> dict:
        h = C.c_double()
        dims = np.zeros(3, dtype=np.uintp)
        padded = np.zeros(3, dtype=np.uintp)
        pe, ne = C.c_size_t(), C.c_size_t()
        _bind_pfft().fho_pfft_stats(self._h, C.byref(h), _ptr(dims), _ptr(padded), C.byref(pe),
                                    C.byref(ne))
        return dict(filaments=self.n, grid_spacing=h.value, grid_dims=[int(d) for d in dims],
                    fft_dims=[int(d) for d in padded], projection_entries=pe.value,
                    near_entries=ne.value)

    def data(self) -> dict:
        st = self.stats()
        pe, ne = st["projection_entries"], st["near_entries"]
        proj_off = np.zeros(self.n + 1, dtype=np.uintp)
        proj_idx = np.zeros(max(pe, 1), dtype=np.uint32)
        proj_w = np.zeros(max(pe, 1))
        near_off = np.zeros(self.n + 1, dtype=np.uintp)
        near_col = np.zeros(max(ne, 1), dtype=np.uint32)
        near_val = np.zeros(max(ne, 1))
        origin = np.zeros(3)
        _bind_pfft().fho_pfft_data(self._h, _ptr(proj_off), _ptr(proj_idx), _ptr(proj_w),
                                   _ptr(near_off), _ptr(near_col), _ptr(near_val), _ptr(origin))
        return dict(proj_off=proj_off.astype(np.int64), proj_idx=proj_idx[:pe],
                    proj_w=proj_w[:pe], near_off=near_off.astype(np.int64),
                    near_col=near_col[:ne], near_val=near_val[:ne], origin=origin, **st)

    def convolve(self, a, b=None):
        a = np.ascontiguousarray(a, dtype=np.float64)
        pa = np.zeros_like(a)
        if b is not None:
            b = np.ascontiguousarray(b, dtype=np.float64)
            pb = np.zeros_like(b)
        _check(_bind_pfft().fho_pfft_convolve(self._h, _ptr(a), _ptr(b) if b is not None else None,
                                              _ptr(pa), _ptr(pb) if b is not None else None))
        return (pa, pb) if b is not None else pa


def gmres_dense(a: np.ndarray, b: np.ndarray, x0=None, tolerance: float = 1e-10,
                restart: int = 100, max_iterations: int = 5000):
    """gmres.rs on a dense matrix without preconditioning: (x, iterations,
    relative_residual, converged)."""
    complex_ = np.iscomplexobj(a) or np.iscomplexobj(b)
    n = len(b)
    A = np.asfortranarray(a, dtype=np.complex128)
    B = np.ascontiguousarray(b, dtype=np.complex128)
    X = np.zeros(n, dtype=np.complex128) if x0 is None else np.array(x0, dtype=np.complex128)
    it, rr, conv = C.c_size_t(), C.c_double(), C.c_int()
    _check(_bind_pfft().fho_gmres_dense(n, _ptr(A), _ptr(B), _ptr(X), tolerance, restart,
                                        max_iterations, 1 if complex_ else 0, C.byref(it),
                                        C.byref(rr), C.byref(conv)))
    return (X if complex_ else X.real.copy()), it.value, rr.value, bool(conv.value)


class IterSystem:
    """IterativeSystem (iterative.rs:155-536) of a PackedProblem."""

    def __init__(self, packed, params: PfftParams = None, tolerance: float = 1e-10,
                 restart: int = 100, max_iterations: int = 5000):
        L = _bind_pfft()
        self.params = params or PfftParams.default()
        self._h = C.c_void_p()
        _check(L.fho_iter_create(_ptr(packed.nodes), len(packed.nodes), _ptr(packed.segments),
                                 len(packed.segments), _ptr(packed.grids), _ptr(packed.ports),
                                 len(packed.ports), C.byref(self.params), tolerance, restart,
                                 max_iterations, C.byref(self._h)))
        c = (C.c_size_t * 6)()
        L.fho_iter_counts(self._h, c)
        (self.nodes, self.segments, self.filaments, self.meshes, self.internal,
         self.ports) = list(c)

    def __del__(self):
        if getattr(self, "_h", None):
            _bind_pfft().fho_iter_destroy(self._h)
            self._h = None

    def impedance(self, frequency: float):
        p = self.ports
        z = np.zeros((p, p), dtype=np.complex128)
        its = np.zeros(max(p, 1), dtype=np.uintp)
        _check(_bind_pfft().fho_iter_impedance(self._h, frequency, _ptr(z), _ptr(its)))
        return z, [int(i) for i in its[:p]]
