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
  // r_mesh, sparse, both triangles (for forming R + jwL)
  const int32_t *d_r_rows = nullptr, *d_r_cols = nullptr;
  const double *d_r_vals = nullptr;
  size_t r_nnz = 0;
  // r_ee as CSR over its m rows (refinement residuals), r_ep dense [q][i]
  const int64_t *d_ree_ptr = nullptr;
  const int32_t *d_ree_col = nullptr;
  const double *d_ree_val = nullptr;
  const double *d_rep = nullptr;
  const double *d_rpp = nullptr;  // r_pp dense p x p
  // r_mesh's lower triangle (row >= column) by column over all m + p rows,
  // rows ascending; may be null (then the matrices are formed in full
  // before the factorization)
  const int64_t *d_rlc_ptr = nullptr;
  const int32_t *d_rlc_row = nullptr;
  const double *d_rlc_val = nullptr;
  // the largest row - column of those entries over the m interior rows (the
  // port rows may hold entries in any column)
  int r_band = 0;
  // refined form only, may be null: receives x = Zee^-1 Zep of every
  // frequency, 2 K p vectors of m (vector 2 (k p + q) the real part of
  // column q at frequency k, the next one its imaginary part)
  double *d_x_out = nullptr;
  // refined form: also refine x to working accuracy before stopping (the
  // default stops when Z, which is stationary in x, has converged)
  bool tight_x = false;
  // refined form: the stages allowed, bit k = stage k (0 split-bf16 factor
  // with classical refinement, 1 the same factor under GMRES-IR, 2 the FP32
  // factor with classical refinement, 3 the FP32 factor under GMRES-IR)
  unsigned refine_stages = 15u;
  // outer block of the factorization (columns), 0 = the form's default
  // (128 FP64, 256 refined); a multiple of 64
  int nb = 0;
  // refined form, may be null: the chunk's graph for replay
  struct RefineGraph *graph = nullptr;
};

// Grow-only device workspace (through device_alloc, so that a batch
// worker's block cache takes it back without a hipFree between problems).
struct LdltWorkspace {
  char *base = nullptr;
  size_t bytes = 0;
  LdltWorkspace() = default;
  LdltWorkspace(const LdltWorkspace &) = delete;
  LdltWorkspace &operator=(const LdltWorkspace &) = delete;
  ~LdltWorkspace() { release(); }
  void release() {
    if (base) device_free(base);
    base = nullptr;
    bytes = 0;
  }
  fh_status reserve(size_t n) {
    if (n <= bytes) return FH_OK;
    release();
    base = static_cast<char *>(device_alloc(n));
    if (!base) return set_error(FH_E_NO_MEMORY, "device workspace of " + std::to_string(n) + " bytes");
    bytes = n;
    return FH_OK;
  }
};

// HIP-event timing of the dominant kernel (the outer trailing update), on the
// launching stream.  Only launches bracketed with begin/end are counted.
struct Profiler {
  std::vector<hipEvent_t> starts, stops;
  size_t used = 0;
  // recording into a captured graph: the events become event-record nodes
  // that stamp each replay
  bool captured = false;
  double flops = 0.0;  // algorithmic (FP32/FP64-equivalent) flops
  double unit_flops = 0.0;  // flops on the matrix-core units used (3x for split bf16)
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
    (void)hipEventRecordWithFlags(starts[used], s, captured ? hipEventRecordExternal : 0);
  }
  void end(hipStream_t s, double f, double unit_factor = 1.0) {
    (void)hipEventRecordWithFlags(stops[used], s, captured ? hipEventRecordExternal : 0);
    ++used;
    flops += f;
    unit_flops += f * unit_factor;
  }
  // Another profiler's intervals (a replayed graph's), added in.
  double offset_ms = 0.0;
  size_t offset_launches = 0;
  void absorb(const Profiler &o) {
    offset_ms += o.total_ms();
    offset_launches += o.used + o.offset_launches;
    flops += o.flops;
    unit_flops += o.unit_flops;
  }
  size_t launches() const { return used + offset_launches; }
  // Sum of the recorded intervals (ms); the stream must be synchronized.
  double total_ms() const {
    double t = offset_ms;
    for (size_t i = 0; i < used; ++i) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, starts[i], stops[i]) == hipSuccess) t += ms;
    }
    return t;
  }
};

// A captured refined solve of one chunk -- the factorization and the first
// refinement steps, with the convergence test on the device -- replayed by
// later sweeps while the launch parameters it captured still hold.  Capture
// starts at the second sweep with the same parameters, so one-off sweeps pay
// nothing for it.
struct RefineGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void *base = nullptr, *omegas = nullptr, *z = nullptr;
  const void *mats[4] = {nullptr, nullptr, nullptr, nullptr};  // L, R_ee, Z_ep, R_scatter
  int K = 0, nb = 0;
  bool profiled = false;
  int seen = 0;
  Profiler prof;  // the events of the captured update launches
  RefineGraph() = default;
  RefineGraph(const RefineGraph &) = delete;
  RefineGraph &operator=(const RefineGraph &) = delete;
  ~RefineGraph() { reset(); }
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    prof.reset_events();
    prof.flops = prof.unit_flops = 0.0;
    seen = 0;
  }
};

// Outcome of a batch solve.
struct BatchResult {
  std::vector<int> singular;         // per frequency: exact zero pivot
  std::vector<int> refine_steps;     // per frequency (refined form)
  std::vector<double> refine_change; // per frequency: final |x^T r| / |b^T x|
  std::vector<int> converged;        // per frequency (refined form)
  std::vector<int> stage;            // per frequency: refinement stage that converged, -1 none
  int factor_arith = 0;              // 0 FP64, 1 FP32 MFMA, 2 FP32 with split-bf16 updates
};

// Plain FP64: Z of K frequencies at once into d_z (K x p x p, {re, im}).
fh_status ldlt_solve_batch(const DenseSolveInputs &in, const double *d_omegas, int K, double *d_z,
                           BatchResult &out, LdltWorkspace &ws, hipStream_t stream,
                           Profiler *prof);

// FP32 factorization on the matrix cores + FP64 iterative refinement of
// x = Zee^-1 Zep, then Z = Zpp - Zep^T x - x^T r (r the final residual).
fh_status ldlt_refined_batch(const DenseSolveInputs &in, const double *d_omegas, int K,
                             double *d_z, BatchResult &out, LdltWorkspace &ws,
                             hipStream_t stream, Profiler *prof);

// Device bytes one frequency of a batch needs (nb: the outer block, 0 =
// the default).
size_t ldlt_bytes_per_frequency(size_t m, size_t p, bool refined, int nb = 0);
// The default outer block of a form.
int ldlt_default_nb(bool refined);

}  // namespace fh
