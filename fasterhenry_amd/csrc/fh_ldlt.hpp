// Per-frequency factorization and Schur complement (fh_ldlt.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "fh_internal.hpp"

namespace fh {

// The frequency-independent mesh data resident on the device.
struct DenseSolveInputs {
  size_t m = 0, p = 0;              // internal loops, ports
  const double *d_lmesh = nullptr;  // (m + p) x (m + p), row-major (symmetric)
  const int32_t *d_r_rows = nullptr, *d_r_cols = nullptr;  // r_mesh, sparse, both triangles
  const double *d_r_vals = nullptr;
  size_t r_nnz = 0;
};

// Grow-only device workspace.
struct LdltWorkspace {
  char *base = nullptr;
  size_t bytes = 0;
  LdltWorkspace() = default;
  LdltWorkspace(const LdltWorkspace &) = delete;
  LdltWorkspace &operator=(const LdltWorkspace &) = delete;
  ~LdltWorkspace() { release(); }
  void release() {
    if (base) (void)hipFree(base);
    base = nullptr;
    bytes = 0;
  }
  fh_status reserve(size_t n) {
    if (n <= bytes) return FH_OK;
    release();
    if (hipMalloc(&base, n) != hipSuccess) {
      base = nullptr;
      (void)hipGetLastError();
      return set_error(FH_E_NO_MEMORY, "device workspace of " + std::to_string(n) + " bytes");
    }
    bytes = n;
    return FH_OK;
  }
};

// HIP-event timing of the dominant kernel family, on the launching stream.
struct Profiler {
  std::vector<hipEvent_t> starts, stops;
  size_t used = 0;
  double flops = 0.0;
  ~Profiler() { reset_events(); }
  void reset_events() {
    for (auto e : starts) (void)hipEventDestroy(e);
    for (auto e : stops) (void)hipEventDestroy(e);
    starts.clear();
    stops.clear();
    used = 0;
  }
  void begin(hipStream_t s) {
    if (used == starts.size()) {
      hipEvent_t a, b;
      (void)hipEventCreate(&a);
      (void)hipEventCreate(&b);
      starts.push_back(a);
      stops.push_back(b);
    }
    (void)hipEventRecord(starts[used], s);
  }
  void end(hipStream_t s, double f) {
    (void)hipEventRecord(stops[used], s);
    ++used;
    flops += f;
  }
  // Sum of the recorded intervals (ms); the stream must be synchronized.
  double total_ms() const {
    double t = 0.0;
    for (size_t i = 0; i < used; ++i) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, starts[i], stops[i]) == hipSuccess) t += ms;
    }
    return t;
  }
};

// FP64 factorization of K frequencies at once; d_z receives K x p x p
// complex values ({re, im}); d_singular[f] is set to 1 on an exact zero
// pivot (it must be cleared by the caller).
fh_status ldlt_solve_batch(const DenseSolveInputs &in, const double *d_omegas, int K, double *d_z,
                           int *d_singular, LdltWorkspace &ws, hipStream_t stream, Profiler *prof);

// Device bytes one frequency of a batch needs.
size_t ldlt_bytes_per_frequency(size_t m, size_t p);

}  // namespace fh
