"""fh_partial_inductance_noalloc: the fill with its scratch in a fixed
per-device array, called through the C ABI on caller-owned device buffers,
against fh_partial_inductance (values and methods bit for bit), on two calls
in a row with the second smaller, and its size limit."""
import ctypes as C

import numpy as np
import pytest

from fasterhenry_amd import fixtures, model, native, partial_inductance_matrix
from fasterhenry_amd.errors import InvalidArgument

pytestmark = pytest.mark.gpu


def _filaments(problem):
    import oracle

    return oracle.discretize(model.pack(*problem))


class _Hip:
    """The HIP runtime the library itself links, for caller-owned device
    buffers and a stream (test plumbing only)."""

    H2D, D2H = 1, 2

    def __init__(self):
        native.lib()  # the same runtime instance the library uses
        self.rt = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        for name in ("hipMalloc", "hipFree", "hipMemcpy", "hipStreamCreate", "hipStreamDestroy",
                     "hipStreamSynchronize"):
            getattr(self.rt, name).restype = C.c_int
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.rt.hipStreamDestroy.argtypes = [C.c_void_p]
        self.rt.hipStreamSynchronize.argtypes = [C.c_void_p]

    def ok(self, rc):
        assert rc == 0, f"HIP runtime call failed: {rc}"

    def upload(self, host: np.ndarray) -> C.c_void_p:
        p = C.c_void_p()
        self.ok(self.rt.hipMalloc(C.byref(p), C.c_size_t(max(host.nbytes, 1))))
        self.ok(self.rt.hipMemcpy(p, host.ctypes.data, host.nbytes, self.H2D))
        return p

    def download(self, p, out: np.ndarray) -> np.ndarray:
        self.ok(self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, self.D2H))
        return out


def _noalloc(f):
    hip = _Hip()
    n = len(f)
    f = np.ascontiguousarray(f)
    d_f = hip.upload(f)
    d_L = hip.upload(np.full(n * n, np.nan))  # poisoned: every entry must be written
    d_m = hip.upload(np.full(n * n, 99, dtype=np.int8))
    stream = C.c_void_p()
    hip.ok(hip.rt.hipStreamCreate(C.byref(stream)))
    try:
        native.check(native.lib().fh_partial_inductance_noalloc(d_f, n, d_L, d_m, stream))
        hip.ok(hip.rt.hipStreamSynchronize(stream))
        L = hip.download(d_L, np.zeros(n * n)).reshape(n, n)
        m = hip.download(d_m, np.zeros(n * n, dtype=np.int8)).reshape(n, n)
    finally:
        hip.ok(hip.rt.hipStreamDestroy(stream))
        for p in (d_f, d_L, d_m):
            hip.ok(hip.rt.hipFree(p))
    return L, m


def test_noalloc_fill_matches_the_allocating_entry_twice_in_a_row():
    first = _filaments(fixtures.random_cloud_problem(700, False, 11))
    second = _filaments(fixtures.busbar_pair(4, 4, 3))
    assert len(second) < len(first)
    for f in (first, second):
        print(f"noalloc fill: {len(f)} filaments", flush=True)
        L, m = _noalloc(f)
        L_ref, m_ref = partial_inductance_matrix(f, with_methods=True)
        assert L.view(np.uint64).tobytes() == L_ref.view(np.uint64).tobytes()
        assert np.array_equal(m, m_ref)
        # the deferred kernel ran: the cloud has queued methods
        assert np.isin(m_ref, [1, 2]).any()


def test_noalloc_fill_rejects_sizes_beyond_its_scratch():
    n = 4001
    lib = native.lib()
    status = lib.fh_partial_inductance_noalloc(C.c_void_p(8), n, C.c_void_p(8), None, None)
    assert status != native.FH_OK
    with pytest.raises(InvalidArgument):
        native.check(status)
    assert native.lib().fh_partial_inductance_noalloc(None, 0, None, None, None) == native.FH_OK
