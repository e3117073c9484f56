// The C ABI (include/fasterhenry_gpu.h): assembled mesh systems resident in
// device memory, frequency sweeps, and the kernel-level entries.
//
// fh_system_create follows MeshSystem::assemble_with_coupling
// (fasterhenry/src/solve.rs:680-766) for the all-pairs coupling; fh_sweep
// follows MeshSystem::sweep / sweep_impedances (solve.rs:881-894,
// 1039-1079): frequencies are validated first, then solved in chunks that
// fit a device-memory budget, and the first failure in list order is
// reported.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <thread>
#include <unordered_map>

#include "fh_adjoint.hpp"
#include "fh_internal.hpp"
#include "fh_iterative.hpp"
#include "fh_ldlt.hpp"
#include "fh_pfft.hpp"

using Clock = std::chrono::steady_clock;

namespace fh {


namespace {

struct BlockCache {
  bool on = false;
  std::multimap<size_t, void *> idle;             // size -> block
  std::unordered_map<void *, size_t> size_of;     // every block this thread's cache handed out
  ~BlockCache() { flush(); }
  void flush() {
    for (auto &kv : idle) (void)hipFree(kv.second);
    idle.clear();
    // blocks still in use go back to plain hipFree when released
    size_of.clear();
  }
};

BlockCache &cache() {
  thread_local BlockCache c;
  return c;
}

}  // namespace

void *device_alloc(size_t bytes) {
  BlockCache &c = cache();
  if (c.on) {
    auto it = c.idle.lower_bound(bytes);
    if (it != c.idle.end() && it->first <= 2 * bytes + (size_t{1} << 20)) {
      void *p = it->second;
      c.idle.erase(it);
      return p;
    }
  }
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    if (!c.on || c.idle.empty()) return nullptr;
    // room may be held by idle blocks
    for (auto &kv : c.idle) {
      c.size_of.erase(kv.second);
      (void)hipFree(kv.second);
    }
    c.idle.clear();
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  if (c.on) c.size_of[p] = bytes;
  return p;
}

void device_free(void *p) {
  BlockCache &c = cache();
  if (c.on) {
    auto it = c.size_of.find(p);
    if (it != c.size_of.end()) {
      c.idle.emplace(it->second, p);
      return;
    }
  }
  (void)hipFree(p);
}

void block_cache(bool on) {
  BlockCache &c = cache();
  if (!on) c.flush();
  c.on = on;
}

}  // namespace fh

using namespace fh;

// Groups of frequencies a refined sweep may factorize side by side.
constexpr int kSweepPartsMax = 4;

struct fh_system {
  hipStream_t stream = nullptr;
  bool own_stream = true;
  fh_solver solver = FH_SOLVER_DENSE;
  std::unique_ptr<IterativeSystem> iter;
  fh_precision precision = FH_FP64_PLAIN;
  bool profiling = false;
  size_t counts[6] = {0, 0, 0, 0, 0, 0};
  std::vector<fh_filament> filaments;
  std::vector<double> resistances;
  LoopBasis basis;
  DeviceArray<fh_filament> d_filaments;
  DeviceArray<double> d_L;
  DeviceArray<double> d_lmesh;
  std::vector<int32_t> r_rows, r_cols;
  std::vector<double> r_vals;
  DeviceArray<int32_t> d_r_rows, d_r_cols;
  DeviceArray<double> d_r_vals;
  // r_ee as CSR (refinement residuals), r_ep dense [q][i], r_pp dense
  DeviceArray<int64_t> d_ree_ptr;
  DeviceArray<int32_t> d_ree_col;
  DeviceArray<double> d_ree_val, d_rep, d_rpp;
  // the lower-triangle entries of r_mesh (row >= column) by column, all
  // N = m + p rows, for forming R + jwL inside the trailing update
  DeviceArray<int64_t> d_rlc_ptr;
  DeviceArray<int32_t> d_rlc_row;
  DeviceArray<double> d_rlc_val;
  int32_t r_band = 0;  // largest row - column of those entries over the m interior rows
  LdltWorkspace ws;
  // more streams and workspaces: a refined sweep chunk is solved as a few
  // parts side by side, so that one part's latency-bound kernels (the
  // diagonal blocks, panels, triangular solves) run under the others' bulk
  // updates
  static constexpr int kParts = kSweepPartsMax;
  int sweep_parts = 2;  // two measured fastest on MI355X (four contend)
  uint32_t refine_stages = FH_REFINE_ALL;
  // outer block of the factorization per form (0: the default), and whether
  // the schedule was set explicitly (then the tuning cache is not consulted)
  int nb_refined = 0, nb_plain = 0;
  bool schedule_fixed = false;
  int nb_for(bool refined) const { return refined ? nb_refined : nb_plain; }
  hipStream_t part_stream[kParts] = {nullptr, nullptr, nullptr, nullptr};
  LdltWorkspace part_ws[kParts];
  // the dense sweep's frequency and Z buffers, kept between sweeps so that a
  // repeated sweep replays each part's captured graph (RefineGraph)
  DeviceArray<double> sweep_omegas, sweep_z;
  RefineGraph part_graph[kParts];
  bool graphs = true;
  double inductance_s = 0.0, assembly_s = 0.0;
  // a copy of the problem (geometry tangents for gradients)
  std::vector<double> prob_nodes;
  std::vector<fh_segment> prob_segments;
  std::vector<fh_grid> prob_grids;
  std::vector<fh_port> prob_ports;
  void keep_problem(const fh_problem &p) {
    prob_nodes.assign(p.nodes, p.nodes + 3 * p.n_nodes);
    prob_segments.assign(p.segments, p.segments + p.n_segments);
    prob_grids.assign(p.grids, p.grids + p.n_segments);
    prob_ports.assign(p.ports, p.ports + p.n_ports);
  }
  fh_problem problem() const {
    fh_problem p;
    p.nodes = prob_nodes.data();
    p.n_nodes = prob_nodes.size() / 3;
    p.segments = prob_segments.data();
    p.grids = prob_grids.data();
    p.n_segments = prob_segments.size();
    p.ports = prob_ports.data();
    p.n_ports = prob_ports.size();
    return p;
  }
  ~fh_system() {
    iter.reset();
    d_filaments.release();
    d_L.release();
    d_lmesh.release();
    ws.release();
    for (int q = 0; q < kParts; ++q) {
      part_ws[q].release();
      if (part_stream[q]) (void)hipStreamDestroy(part_stream[q]);
    }
    if (stream && own_stream) (void)hipStreamDestroy(stream);
  }
};

struct fh_pfft {
  hipStream_t stream = nullptr;
  PfftOperator op;
  DeviceArray<double> dx, dy;
  ~fh_pfft() {
    dx.release();
    dy.release();
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// A batch worker's stream, lent to the systems it creates (none elsewhere).
thread_local hipStream_t tl_worker_stream = nullptr;

fh_status system_stream(fh_system &s) {
  if (tl_worker_stream) {
    s.stream = tl_worker_stream;
    s.own_stream = false;
    return FH_OK;
  }
  FH_HIP_TRY(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
  return FH_OK;
}

double seconds_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

fh_status check_device() {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    (void)hipGetLastError();
    return set_error(FH_E_HIP, "no HIP device is visible");
  }
  return FH_OK;
}

// Frequencies per part of a split refined sweep chunk, at least.
constexpr int kSplitMin = 4;

size_t sweep_budget_bytes() {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
    (void)hipGetLastError();
    return size_t(8) << 30;
  }
  // Leave headroom; never ask for more than 96 GiB of workspace.
  size_t b = free_b / 2;
  return std::min(b, size_t(96) << 30);
}


// A host pair mask (group per filament, symmetric keep table) to the device.
fh_status upload_mask(const fh_pair_mask *mask, DeviceArray<int32_t> &d_group, DeviceArray<uint8_t> &d_table,
                      DevicePairMask &out, hipStream_t stream) {
  out = DevicePairMask{};
  if (!mask || mask->n == 0) return FH_OK;
  fh_status st;
  const size_t tab = mask->groups * mask->groups;
  if ((st = d_group.allocate(mask->n)) || (st = d_table.allocate(std::max<size_t>(tab, 1)))) return st;
  FH_HIP_TRY(hipMemcpyAsync(d_group.ptr, mask->group, mask->n * sizeof(int32_t), hipMemcpyHostToDevice, stream));
  if (tab)
    FH_HIP_TRY(hipMemcpyAsync(d_table.ptr, mask->coupled, tab, hipMemcpyHostToDevice, stream));
  out.group = d_group.ptr;
  out.n = mask->n;
  out.groups = mask->groups;
  out.table = d_table.ptr;
  return FH_OK;
}

// The tuning cache's schedule for this system's size class, unless the
// caller fixed one (fh_system_set_schedule).
void apply_tuned_schedule(fh_system *s, bool refined, size_t K) {
  if (s->schedule_fixed) return;
  TunedSchedule t;
  if (!tune_lookup(dense_tune_key(refined, s->basis.internal, K), t)) return;
  if (refined) {
    s->nb_refined = t.outer_block;
    if (t.sweep_parts >= 1 && t.sweep_parts <= kSweepPartsMax) s->sweep_parts = t.sweep_parts;
  } else {
    s->nb_plain = t.outer_block;
  }
}

// The frequencies of a refined chunk that no FP32 stage accepted, solved by
// the plain FP64 form (ldlt_solve_batch) in groups that fit the memory
// budget; their Z goes to z_chunk (kc x p x p {re, im}, host) and res marks
// them converged at stage 4 (singular flags from the FP64 pivots).
fh_status fp64_resolve(fh_system *s, const DenseSolveInputs &in, const double *omegas, size_t kc,
                       BatchResult &res, double *z_chunk) {
  std::vector<size_t> left;
  for (size_t q = 0; q < kc; ++q)
    if (!res.converged[q]) left.push_back(q);
  if (left.empty()) return FH_OK;
  const size_t p = in.p;
  DenseSolveInputs in64 = in;
  in64.nb = s->nb_plain;
  const size_t per = ldlt_bytes_per_frequency(in.m, p, false, in64.nb);
  const size_t group = std::min(left.size(), std::max<size_t>(1, sweep_budget_bytes() / per));
  DeviceArray<double> d_w, d_z;
  fh_status st;
  if ((st = d_w.allocate(group)) || (st = d_z.allocate(group * p * p * 2))) return st;
  LdltWorkspace ws;
  std::vector<double> w(group), z(group * p * p * 2);
  for (size_t g0 = 0; g0 < left.size(); g0 += group) {
    const size_t gn = std::min(group, left.size() - g0);
    for (size_t k = 0; k < gn; ++k) w[k] = omegas[left[g0 + k]];
    FH_HIP_TRY(hipMemcpyAsync(d_w.ptr, w.data(), gn * sizeof(double), hipMemcpyHostToDevice, s->stream));
    BatchResult r64;
    if ((st = ldlt_solve_batch(in64, d_w.ptr, static_cast<int>(gn), d_z.ptr, r64, ws, s->stream, nullptr)))
      return st;
    FH_HIP_TRY(hipMemcpyAsync(z.data(), d_z.ptr, gn * p * p * 2 * sizeof(double), hipMemcpyDeviceToHost,
                              s->stream));
    FH_HIP_TRY(hipStreamSynchronize(s->stream));
    for (size_t k = 0; k < gn; ++k) {
      const size_t q = left[g0 + k];
      std::memcpy(z_chunk + 2 * p * p * q, z.data() + 2 * p * p * k, 2 * p * p * sizeof(double));
      res.singular[q] = r64.singular[k];
      res.converged[q] = 1;
      res.stage[q] = 4;
    }
  }
  return FH_OK;
}

// x = Zee^-1 Zep (2 K p vectors of m: real then imaginary part of column a
// at frequency k, vector 2 (k p + a)) by the refined solve, refined until
// the correction is at the rounding level (tight_x); first failure in list
// order as in fh_sweep.  The system must be dense; nothing to solve for m = 0.
fh_status refined_x(fh_system *s, const double *freqs, size_t K, std::vector<double> &X) {
  const size_t m = s->basis.internal, p = s->basis.loops - m;
  X.assign(2 * K * p * m, 0.0);
  if (m == 0 || K == 0) return FH_OK;
  const size_t per = ldlt_bytes_per_frequency(m, p, true, s->nb_refined) + 2 * p * m * sizeof(double);
  const size_t chunk = std::min(K, std::max<size_t>(1, sweep_budget_bytes() / per));
  DeviceArray<double> d_omegas, d_z, d_x;
  fh_status st;
  if ((st = d_omegas.allocate(chunk)) || (st = d_z.allocate(chunk * p * p * 2)) ||
      (st = d_x.allocate(2 * chunk * p * m)))
    return st;
  DenseSolveInputs in;
  in.m = m;
  in.p = p;
  in.d_lmesh = s->d_lmesh.ptr;
  in.d_r_rows = s->d_r_rows.ptr;
  in.d_r_cols = s->d_r_cols.ptr;
  in.d_r_vals = s->d_r_vals.ptr;
  in.r_nnz = s->r_vals.size();
  in.d_ree_ptr = s->d_ree_ptr.ptr;
  in.d_ree_col = s->d_ree_col.ptr;
  in.d_ree_val = s->d_ree_val.ptr;
  in.d_rep = s->d_rep.ptr;
  in.d_rpp = s->d_rpp.ptr;
  in.d_rlc_ptr = s->d_rlc_ptr.ptr;
  in.d_rlc_row = s->d_rlc_row.ptr;
  in.d_rlc_val = s->d_rlc_val.ptr;
  in.r_band = s->r_band;
  in.d_x_out = d_x.ptr;
  in.tight_x = true;
  in.refine_stages = s->refine_stages & 15u;
  in.nb = s->nb_refined;
  std::vector<double> omegas(chunk), z(2 * chunk * p * p);
  for (size_t c0 = 0; c0 < K; c0 += chunk) {
    const size_t kc = std::min(chunk, K - c0);
    for (size_t q = 0; q < kc; ++q) omegas[q] = 6.283185307179586 * freqs[c0 + q];
    FH_HIP_TRY(hipMemcpyAsync(d_omegas.ptr, omegas.data(), kc * sizeof(double), hipMemcpyHostToDevice, s->stream));
    BatchResult res;
    if ((st = ldlt_refined_batch(in, d_omegas.ptr, static_cast<int>(kc), d_z.ptr, res, s->ws, s->stream, nullptr)))
      return st;
    FH_HIP_TRY(hipMemcpyAsync(z.data(), d_z.ptr, kc * p * p * 2 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    FH_HIP_TRY(hipMemcpyAsync(X.data() + 2 * p * m * c0, d_x.ptr, 2 * kc * p * m * sizeof(double),
                              hipMemcpyDeviceToHost, s->stream));
    FH_HIP_TRY(hipStreamSynchronize(s->stream));
    for (size_t q = 0; q < kc; ++q) {
      bool bad = res.singular[q] != 0;
      for (size_t e = 0; e < 2 * p * p && !bad; ++e)
        if (!std::isfinite(z[2 * q * p * p + e])) bad = true;
      for (size_t e = 0; e < 2 * p * m && !bad; ++e)
        if (!std::isfinite(X[2 * p * m * (c0 + q) + e])) bad = true;
      if (bad)
        return set_error(FH_E_SINGULAR,
                         "the mesh system is singular or ill-conditioned at " + std::to_string(freqs[c0 + q]) + " Hz",
                         static_cast<int64_t>(c0 + q), 0, freqs[c0 + q]);
      if (!res.converged[q])
        return set_error(FH_E_NOT_CONVERGED,
                         "iterative refinement did not converge at " + std::to_string(freqs[c0 + q]) + " Hz",
                         static_cast<int64_t>(c0 + q), res.refine_steps[q], freqs[c0 + q]);
    }
  }
  return FH_OK;
}

}  // namespace

extern "C" {

fh_solver fh_resolve(fh_solver_choice choice, size_t n) {
  if (choice == FH_DENSE) return FH_SOLVER_DENSE;
  if (choice == FH_ITERATIVE) return FH_SOLVER_ITERATIVE;
  return n <= FH_DENSE_PATH_MAX_FILAMENTS ? FH_SOLVER_DENSE : FH_SOLVER_ITERATIVE;
}


fh_status fh_system_create(const fh_problem *problem, fh_solver_choice choice,
                           fh_precision precision, fh_system **out) {
  return fh_system_create_coupled(problem, choice, precision, nullptr, out);
}

fh_status fh_system_create_coupled(const fh_problem *problem, fh_solver_choice choice, fh_precision precision,
                                   const fh_coupling *coupling, fh_system **out) {
  clear_error();
  if (!problem || !out) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (precision != FH_FP64_PLAIN && precision != FH_FP32_REFINED)
    return set_error(FH_E_INVALID_ARGUMENT, "unknown precision");
  const auto t0 = Clock::now();
  size_t nfil = 0;
  for (size_t s = 0; s < problem->n_segments && problem->grids; ++s)
    nfil += static_cast<size_t>(problem->grids[s].nw) * problem->grids[s].nh;
  if (choice != FH_AUTO && choice != FH_DENSE && choice != FH_ITERATIVE)
    return set_error(FH_E_INVALID_ARGUMENT, "unknown solver choice");
  const bool truncating = coupling && coupling->coupled;
  if (truncating) {
    if (coupling->n_segments != problem->n_segments || (problem->n_segments && !coupling->segment_group))
      return set_error(FH_E_INVALID_ARGUMENT, "the coupling must give one group per segment");
    // fasterhenry-cli/src/lib.rs:89-107: the matrix-free operator couples
    // every pair, so a truncation stays dense
    if (choice == FH_ITERATIVE)
      return set_error(FH_E_UNSUPPORTED,
                       "the iterative solver cannot honour a coupling truncation: the matrix-free operator "
                       "couples every pair of filaments");
  }
  if (!truncating && fh_resolve(choice, nfil) != FH_SOLVER_DENSE) {
    fh_iterative_params params;
    fh_iterative_default_params(&params);
    fh_status st = fh_system_create_iterative(problem, &params, out);
    if (st == FH_OK) (*out)->precision = precision;
    return st;
  }
  fh_status st = check_device();
  if (st) return st;
  std::unique_ptr<fh_system> s(new fh_system());
  s->precision = precision;
  if ((st = discretize_problem(*problem, s->filaments))) return st;
  s->keep_problem(*problem);
  const size_t n = s->filaments.size();
  std::vector<size_t> per(problem->n_segments);
  for (size_t q = 0; q < problem->n_segments; ++q)
    per[q] = static_cast<size_t>(problem->grids[q].nw) * problem->grids[q].nh;
  if ((st = build_loop_basis(*problem, per, s->basis))) return st;
  s->resistances.resize(n);
  for (size_t i = 0; i < n; ++i) {
    const fh_filament &f = s->filaments[i];
    s->resistances[i] = f.length / (f.sigma * (f.width * f.height));  // solve.rs:941-943
  }
  if ((st = system_stream(*s))) return st;
  if ((st = s->d_filaments.allocate(n))) return st;
  if ((st = s->d_L.allocate(n * n))) return st;
  if (n) {
    FH_HIP_TRY(hipMemcpyAsync(s->d_filaments.ptr, s->filaments.data(), n * sizeof(fh_filament),
                              hipMemcpyHostToDevice, s->stream));
  }
  // PairMask from the segment groups (Coupling::pair_mask, coupling.rs:259-272)
  std::vector<int32_t> fgroup;
  fh_pair_mask hmask{nullptr, 0, 0, nullptr};
  DeviceArray<int32_t> d_g;
  DeviceArray<uint8_t> d_t;
  DevicePairMask dmask;
  if (truncating) {
    fgroup.reserve(n);
    for (size_t q = 0; q < problem->n_segments; ++q) fgroup.insert(fgroup.end(), per[q], coupling->segment_group[q]);
    hmask = fh_pair_mask{fgroup.data(), fgroup.size(), coupling->groups, coupling->coupled};
    if ((st = upload_mask(&hmask, d_g, d_t, dmask, s->stream))) return st;
  }
  const auto ti = Clock::now();
  if ((st = fill_inductance(s->d_filaments.ptr, n, s->d_L.ptr, nullptr, s->stream, truncating ? &dmask : nullptr)))
    return st;
  s->inductance_s = seconds_since(ti);
  const size_t loops = s->basis.loops;
  if ((st = s->d_lmesh.allocate(loops * loops))) return st;
  if ((st = congruence_inductance(s->basis, n, s->d_L.ptr, s->d_lmesh.ptr, s->stream))) return st;
  congruence_resistance(s->basis, s->resistances, s->r_rows, s->r_cols, s->r_vals);
  const size_t nnz = s->r_vals.size();
  if ((st = s->d_r_rows.allocate(nnz)) || (st = s->d_r_cols.allocate(nnz)) ||
      (st = s->d_r_vals.allocate(nnz)))
    return st;
  if (nnz) {
    FH_HIP_TRY(hipMemcpyAsync(s->d_r_rows.ptr, s->r_rows.data(), nnz * sizeof(int32_t),
                              hipMemcpyHostToDevice, s->stream));
    FH_HIP_TRY(hipMemcpyAsync(s->d_r_cols.ptr, s->r_cols.data(), nnz * sizeof(int32_t),
                              hipMemcpyHostToDevice, s->stream));
    FH_HIP_TRY(hipMemcpyAsync(s->d_r_vals.ptr, s->r_vals.data(), nnz * sizeof(double),
                              hipMemcpyHostToDevice, s->stream));
  }
  {
    // CSR of the r_ee block, dense r_ep ([q][i]) and r_pp, for the refined form.
    const size_t m = s->basis.internal, p = loops - m;
    std::vector<std::vector<std::pair<int32_t, double>>> rows(m);
    std::vector<double> rep(m * p, 0.0), rpp(p * p, 0.0);
    for (size_t e = 0; e < nnz; ++e) {
      const size_t i = static_cast<size_t>(s->r_rows[e]), j = static_cast<size_t>(s->r_cols[e]);
      if (i < m && j < m) rows[i].emplace_back(static_cast<int32_t>(j), s->r_vals[e]);
      else if (i < m && j >= m) rep[(j - m) * m + i] = s->r_vals[e];
      else if (i >= m && j >= m) rpp[(i - m) * p + (j - m)] = s->r_vals[e];
    }
    std::vector<int64_t> ptr(m + 1, 0);
    std::vector<int32_t> col;
    std::vector<double> val;
    for (size_t i = 0; i < m; ++i) {
      std::sort(rows[i].begin(), rows[i].end());
      for (auto &e : rows[i]) {
        col.push_back(e.first);
        val.push_back(e.second);
      }
      ptr[i + 1] = static_cast<int64_t>(col.size());
    }
    if ((st = s->d_ree_ptr.allocate(m + 1)) || (st = s->d_ree_col.allocate(std::max<size_t>(col.size(), 1))) ||
        (st = s->d_ree_val.allocate(std::max<size_t>(val.size(), 1))) ||
        (st = s->d_rep.allocate(std::max<size_t>(rep.size(), 1))) ||
        (st = s->d_rpp.allocate(std::max<size_t>(rpp.size(), 1))))
      return st;
    FH_HIP_TRY(hipMemcpyAsync(s->d_ree_ptr.ptr, ptr.data(), ptr.size() * sizeof(int64_t),
                              hipMemcpyHostToDevice, s->stream));
    if (!col.empty()) {
      FH_HIP_TRY(hipMemcpyAsync(s->d_ree_col.ptr, col.data(), col.size() * sizeof(int32_t),
                                hipMemcpyHostToDevice, s->stream));
      FH_HIP_TRY(hipMemcpyAsync(s->d_ree_val.ptr, val.data(), val.size() * sizeof(double),
                                hipMemcpyHostToDevice, s->stream));
    }
    if (!rep.empty())
      FH_HIP_TRY(hipMemcpyAsync(s->d_rep.ptr, rep.data(), rep.size() * sizeof(double),
                                hipMemcpyHostToDevice, s->stream));
    if (!rpp.empty())
      FH_HIP_TRY(hipMemcpyAsync(s->d_rpp.ptr, rpp.data(), rpp.size() * sizeof(double),
                                hipMemcpyHostToDevice, s->stream));
    // r_mesh's lower triangle by column (rows ascending): the entries the
    // scatter writes (row >= column), for the update that forms them
    std::vector<std::vector<std::pair<int32_t, double>>> cols(loops);
    for (size_t e = 0; e < nnz; ++e)
      if (s->r_rows[e] >= s->r_cols[e])
        cols[static_cast<size_t>(s->r_cols[e])].emplace_back(s->r_rows[e], s->r_vals[e]);
    std::vector<int64_t> cptr(loops + 1, 0);
    std::vector<int32_t> crow;
    std::vector<double> cval;
    s->r_band = 0;
    for (size_t e = 0; e < nnz; ++e)
      if (s->r_rows[e] >= s->r_cols[e] && static_cast<size_t>(s->r_rows[e]) < m)
        s->r_band = std::max(s->r_band, s->r_rows[e] - s->r_cols[e]);
    for (size_t c = 0; c < loops; ++c) {
      std::sort(cols[c].begin(), cols[c].end());
      for (auto &e : cols[c]) {
        crow.push_back(e.first);
        cval.push_back(e.second);
      }
      cptr[c + 1] = static_cast<int64_t>(crow.size());
    }
    if ((st = s->d_rlc_ptr.allocate(loops + 1)) ||
        (st = s->d_rlc_row.allocate(std::max<size_t>(crow.size(), 1))) ||
        (st = s->d_rlc_val.allocate(std::max<size_t>(cval.size(), 1))))
      return st;
    FH_HIP_TRY(hipMemcpyAsync(s->d_rlc_ptr.ptr, cptr.data(), cptr.size() * sizeof(int64_t),
                              hipMemcpyHostToDevice, s->stream));
    if (!crow.empty()) {
      FH_HIP_TRY(hipMemcpyAsync(s->d_rlc_row.ptr, crow.data(), crow.size() * sizeof(int32_t),
                                hipMemcpyHostToDevice, s->stream));
      FH_HIP_TRY(hipMemcpyAsync(s->d_rlc_val.ptr, cval.data(), cval.size() * sizeof(double),
                                hipMemcpyHostToDevice, s->stream));
    }
    FH_HIP_TRY(hipStreamSynchronize(s->stream));
  }
  s->counts[0] = problem->n_nodes;
  s->counts[1] = problem->n_segments;
  s->counts[2] = n;
  s->counts[3] = loops;
  s->counts[4] = s->basis.internal;
  s->counts[5] = loops - s->basis.internal;
  s->assembly_s = seconds_since(t0);
  *out = s.release();
  return FH_OK;
}

void fh_iterative_default_params(fh_iterative_params *out) {
  if (!out) return;
  fh_pfft_default_params(&out->pfft);
  out->gmres.tolerance = 1e-10;      // gmres.rs:45
  out->gmres.restart = 100;          // gmres.rs:47
  out->gmres.max_iterations = 5000;  // gmres.rs:49
}

void fh_pfft_default_params(fh_pfft_params *out) {
  if (!out) return;
  out->auto_grid = 1;
  out->interpolation_order = 5;   // pfft/mod.rs:217
  out->cells_per_filament = 8.0;  // pfft/mod.rs:213
  out->spacing = 0.0;
  out->near_field_radius = 8.0;   // pfft/mod.rs:215
  out->fft_precision = 0;
  out->reserved = 0;
}

fh_status fh_system_create_iterative(const fh_problem *problem, const fh_iterative_params *params,
                                     fh_system **out) {
  clear_error();
  if (!problem || !params || !out) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  const auto t0 = Clock::now();
  fh_status st = check_device();
  if (st) return st;
  std::unique_ptr<fh_system> s(new fh_system());
  s->solver = FH_SOLVER_ITERATIVE;
  if ((st = discretize_problem(*problem, s->filaments))) return st;
  s->keep_problem(*problem);
  const size_t n = s->filaments.size();
  std::vector<size_t> per(problem->n_segments);
  for (size_t q = 0; q < problem->n_segments; ++q)
    per[q] = static_cast<size_t>(problem->grids[q].nw) * problem->grids[q].nh;
  if ((st = build_loop_basis(*problem, per, s->basis))) return st;
  s->resistances.resize(n);
  for (size_t i = 0; i < n; ++i) {
    const fh_filament &f = s->filaments[i];
    s->resistances[i] = f.length / (f.sigma * (f.width * f.height));  // solve.rs:941-943
  }
  if ((st = system_stream(*s))) return st;
  s->iter.reset(new IterativeSystem());
  if ((st = s->iter->build(s->filaments, s->resistances, s->basis, *params, s->stream))) return st;
  s->inductance_s = s->iter->operator_seconds();
  const size_t loops = s->basis.loops;
  s->counts[0] = problem->n_nodes;
  s->counts[1] = problem->n_segments;
  s->counts[2] = n;
  s->counts[3] = loops;
  s->counts[4] = s->basis.internal;
  s->counts[5] = loops - s->basis.internal;
  s->assembly_s = seconds_since(t0);
  *out = s.release();
  return FH_OK;
}

fh_status fh_system_operator_stats(const fh_system *s, fh_pfft_stats *out) {
  clear_error();
  if (!s || !out) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  if (!s->iter) return set_error(FH_E_UNSUPPORTED, "a dense system has no pFFT operator");
  *out = s->iter->op().stats();
  return FH_OK;
}

void fh_system_destroy(fh_system *system) { delete system; }

fh_status fh_system_set_precision(fh_system *s, fh_precision precision) {
  clear_error();
  if (!s) return set_error(FH_E_INVALID_ARGUMENT, "null system");
  if (precision != FH_FP64_PLAIN && precision != FH_FP32_REFINED)
    return set_error(FH_E_INVALID_ARGUMENT, "unknown precision");
  s->precision = precision;
  return FH_OK;
}

fh_status fh_system_set_profiling(fh_system *s, int enable) {
  clear_error();
  if (!s) return set_error(FH_E_INVALID_ARGUMENT, "null system");
  s->profiling = enable != 0;
  return FH_OK;
}

fh_status fh_system_counts(const fh_system *s, size_t counts[6]) {
  clear_error();
  if (!s || !counts) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  std::memcpy(counts, s->counts, sizeof(s->counts));
  return FH_OK;
}

fh_status fh_system_filaments(const fh_system *s, fh_filament *out) {
  clear_error();
  if (!s || !out) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  std::memcpy(out, s->filaments.data(), s->filaments.size() * sizeof(fh_filament));
  return FH_OK;
}

fh_status fh_system_inductance(const fh_system *s, double *L_out) {
  clear_error();
  if (!s || !L_out) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  if (s->iter) return set_error(FH_E_UNSUPPORTED, "an iterative system holds no dense L");
  const size_t n = s->filaments.size();
  if (n) FH_HIP_TRY(hipMemcpy(L_out, s->d_L.ptr, n * n * sizeof(double), hipMemcpyDeviceToHost));
  return FH_OK;
}

fh_status fh_system_blocks(const fh_system *s, double *r_ee, double *l_ee, double *r_ep,
                           double *l_ep, double *r_pp, double *l_pp) {
  clear_error();
  if (!s) return set_error(FH_E_INVALID_ARGUMENT, "null system");
  if (s->iter) return set_error(FH_E_UNSUPPORTED, "an iterative system holds no dense mesh blocks");
  const size_t loops = s->basis.loops, m = s->basis.internal, p = loops - m;
  std::vector<double> lm(loops * loops), rm(loops * loops, 0.0);
  if (loops)
    FH_HIP_TRY(hipMemcpy(lm.data(), s->d_lmesh.ptr, lm.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (size_t e = 0; e < s->r_vals.size(); ++e)
    rm[static_cast<size_t>(s->r_rows[e]) * loops + s->r_cols[e]] = s->r_vals[e];
  for (size_t i = 0; i < loops; ++i)
    for (size_t j = 0; j < loops; ++j) {
      const double r = rm[i * loops + j], l = lm[i * loops + j];
      if (i < m && j < m) {
        if (r_ee) r_ee[i * m + j] = r;
        if (l_ee) l_ee[i * m + j] = l;
      } else if (i < m) {
        if (r_ep) r_ep[i * p + (j - m)] = r;
        if (l_ep) l_ep[i * p + (j - m)] = l;
      } else if (j >= m) {
        if (r_pp) r_pp[(i - m) * p + (j - m)] = r;
        if (l_pp) l_pp[(i - m) * p + (j - m)] = l;
      }
    }
  return FH_OK;
}

fh_status fh_sweep(fh_system *s, const double *freqs, size_t K, double *z_out,
                   fh_sweep_info *info) {
  clear_error();
  if (!s || (K && (!freqs || !z_out))) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  for (size_t k = 0; k < K; ++k)
    if (!(std::isfinite(freqs[k]) && freqs[k] >= 0.0))
      return set_error(FH_E_INVALID_FREQUENCY,
                       "frequency " + std::to_string(k) + " is " + std::to_string(freqs[k]) +
                           " Hz; frequencies must be finite and non-negative",
                       static_cast<int64_t>(k), 0, freqs[k]);
  const auto t0 = Clock::now();
  const size_t m = s->basis.internal, p = s->basis.loops - m;
  if (s->iter) {
    // IterativeSystem::sweep (iterative.rs:387-404): frequencies in list
    // order, the first failure reported with its index.
    IterativeSystem &it = *s->iter;
    it.reset_counters();
    it.set_profiling(s->profiling);
    int32_t it_max = 0;
    int64_t it_total = 0;
    double res_max = 0.0;
    std::vector<size_t> iters;
    for (size_t k = 0; k < K; ++k) {
      fh_status st = it.impedance(freqs[k], z_out + 2 * p * p * k, iters, res_max);
      if (st == FH_E_NOT_CONVERGED) {
        ErrorState &e = error_state();  // index1 = port, value = residual, code = iterations
        e.index0 = static_cast<int64_t>(k);
        e.message += " (frequency index " + std::to_string(k) + ")";
        return st;
      }
      if (st) return st;
      for (size_t q = 0; q < iters.size(); ++q) {
        it_max = std::max<int32_t>(it_max, static_cast<int32_t>(iters[q]));
        it_total += static_cast<int64_t>(iters[q]);
      }
      for (size_t e = 0; e < 2 * p * p; ++e)
        if (!std::isfinite(z_out[2 * p * p * k + e]))
          return set_error(FH_E_SINGULAR,
                           "the mesh system is singular or ill-conditioned at " + std::to_string(freqs[k]) + " Hz",
                           static_cast<int64_t>(k), 0, freqs[k]);
    }
    if (info) {
      std::memset(info, 0, sizeof(*info));
      info->nodes = s->counts[0];
      info->segments = s->counts[1];
      info->filaments = s->counts[2];
      info->meshes = s->counts[3];
      info->internal_meshes = s->counts[4];
      info->ports = s->counts[5];
      info->inductance_s = s->inductance_s;
      info->assembly_s = s->assembly_s;
      info->solve_s = seconds_since(t0);
      info->solver = FH_SOLVER_ITERATIVE;
      info->gmres_iterations_max = it_max;
      info->gmres_iterations = it_total;
      info->gmres_residual_max = res_max;
      info->operator_applies = static_cast<int64_t>(it.applies());
      info->operator_apply_ms = s->profiling ? it.apply_ms() : 0.0;
      info->refine_stage = -1;
    }
    return FH_OK;
  }
  Profiler prof;
  Profiler *pp = s->profiling ? &prof : nullptr;
  double extra_ms = 0.0, extra_flops = 0.0, extra_unit_flops = 0.0;  // the parts' profiled launches
  size_t extra_launches = 0;
  int refine_steps_max = 0;
  double refine_change_max = 0.0;
  int factor_arith = s->precision == FH_FP32_REFINED ? 1 : 0;
  int refine_stage = -1;
  if (m == 0) {
    // No internal loops: Z = Zpp, nothing to solve (solve.rs:1159-1161).
    std::vector<double> lm(p * p);
    if (p)
      FH_HIP_TRY(hipMemcpy(lm.data(), s->d_lmesh.ptr, lm.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<double> rm(p * p, 0.0);
    for (size_t e = 0; e < s->r_vals.size(); ++e) rm[s->r_rows[e] * p + s->r_cols[e]] = s->r_vals[e];
    for (size_t k = 0; k < K; ++k) {
      const double omega = 6.283185307179586 * freqs[k];
      for (size_t q = 0; q < p * p; ++q) {
        z_out[2 * (k * p * p + q)] = rm[q];
        z_out[2 * (k * p * p + q) + 1] = freqs[k] == 0.0 ? 0.0 : omega * lm[q];
      }
    }
  } else {
    const bool refined = s->precision == FH_FP32_REFINED;
    apply_tuned_schedule(s, refined, K);
    const size_t per = ldlt_bytes_per_frequency(m, p, refined, s->nb_for(refined));
    size_t chunk = std::max<size_t>(1, sweep_budget_bytes() / per);
    chunk = std::min(chunk, K);
    DeviceArray<double> &d_omegas = s->sweep_omegas, &d_z = s->sweep_z;
    fh_status st;
    if (d_omegas.count < std::max<size_t>(chunk, 1) &&
        (st = d_omegas.allocate(std::max<size_t>(chunk, 1))))
      return st;
    if (d_z.count < std::max<size_t>(chunk * p * p * 2, 1) &&
        (st = d_z.allocate(std::max<size_t>(chunk * p * p * 2, 1))))
      return st;
    DenseSolveInputs in;
    in.m = m;
    in.p = p;
    in.d_lmesh = s->d_lmesh.ptr;
    in.d_r_rows = s->d_r_rows.ptr;
    in.d_r_cols = s->d_r_cols.ptr;
    in.d_r_vals = s->d_r_vals.ptr;
    in.r_nnz = s->r_vals.size();
    in.d_ree_ptr = s->d_ree_ptr.ptr;
    in.d_ree_col = s->d_ree_col.ptr;
    in.d_ree_val = s->d_ree_val.ptr;
    in.d_rep = s->d_rep.ptr;
    in.d_rpp = s->d_rpp.ptr;
    in.d_rlc_ptr = s->d_rlc_ptr.ptr;
    in.d_rlc_row = s->d_rlc_row.ptr;
    in.d_rlc_val = s->d_rlc_val.ptr;
    in.r_band = s->r_band;
    in.refine_stages = s->refine_stages & 15u;
    in.nb = s->nb_for(refined);
    std::vector<double> omegas(chunk);
    for (size_t c0 = 0; c0 < K; c0 += chunk) {
      const size_t kc = std::min(chunk, K - c0);
      for (size_t q = 0; q < kc; ++q) omegas[q] = 6.283185307179586 * freqs[c0 + q];
      FH_HIP_TRY(hipMemcpyAsync(d_omegas.ptr, omegas.data(), kc * sizeof(double),
                                hipMemcpyHostToDevice, s->stream));
      BatchResult res;
      // parts side by side: a refined chunk of enough frequencies on a
      // system large enough for the parts to fill the device, outside the
      // batch workers (which already run one system per stream)
      const int want = s->sweep_parts;
      const int parts = !(refined && m >= 1024 && !tl_worker_stream && s->own_stream) ? 1
                        : std::max(1, std::min<int>(want, static_cast<int>(kc / kSplitMin)));
      if (parts > 1) {
        FH_HIP_TRY(hipStreamSynchronize(s->stream));  // the frequencies are uploaded
        int dev = 0;
        FH_HIP_TRY(hipGetDevice(&dev));
        std::vector<size_t> off(parts + 1, 0);
        for (int q = 0; q < parts; ++q) off[q + 1] = off[q] + (kc - off[q]) / (parts - q);
        std::vector<BatchResult> rs(parts);
        std::vector<Profiler> profs(parts);
        std::vector<fh_status> sts(parts, FH_OK);
        std::vector<ErrorState> errs(parts);
        std::vector<double> ms(parts, 0.0);
        std::vector<size_t> launches(parts, 0);
        for (int q = 1; q < parts; ++q)
          if (!s->part_stream[q]) FH_HIP_TRY(hipStreamCreateWithFlags(&s->part_stream[q], hipStreamNonBlocking));
        auto run_part = [&](int q) {
          hipStream_t sq = q == 0 ? s->stream : s->part_stream[q];
          LdltWorkspace &wq = q == 0 ? s->ws : s->part_ws[q];
          const size_t k0 = off[q], kn = off[q + 1] - off[q];
          DenseSolveInputs inq = in;
          inq.graph = s->graphs ? &s->part_graph[q] : nullptr;
          fh_status stq = q == 0 || hipSetDevice(dev) == hipSuccess
                              ? ldlt_refined_batch(inq, d_omegas.ptr + k0, static_cast<int>(kn), d_z.ptr + 2 * p * p * k0,
                                                   rs[q], wq, sq, pp ? &profs[q] : nullptr)
                              : set_error(FH_E_HIP, "hipSetDevice failed in a sweep part");
          if (!stq && hipStreamSynchronize(sq) != hipSuccess)
            stq = set_error(FH_E_HIP, "hipStreamSynchronize failed on a sweep part's stream");
          if (!stq && pp) {
            ms[q] = profs[q].total_ms();
            launches[q] = profs[q].launches();
          }
          if (stq) errs[q] = error_state();
          sts[q] = stq;
        };
        std::vector<std::thread> workers;
        std::vector<int> here;  // parts whose thread could not be started
        for (int q = 1; q < parts; ++q) {
          try {
            workers.emplace_back(run_part, q);
          } catch (...) {
            here.push_back(q);
          }
        }
        run_part(0);
        for (const int q : here) run_part(q);
        for (auto &t : workers) t.join();
        for (int q = 0; q < parts; ++q)
          if (sts[q]) {
            error_state() = errs[q];
            return sts[q];
          }
        res = BatchResult{};
        for (int q = 0; q < parts; ++q) {
          res.singular.insert(res.singular.end(), rs[q].singular.begin(), rs[q].singular.end());
          res.refine_steps.insert(res.refine_steps.end(), rs[q].refine_steps.begin(), rs[q].refine_steps.end());
          res.refine_change.insert(res.refine_change.end(), rs[q].refine_change.begin(), rs[q].refine_change.end());
          res.converged.insert(res.converged.end(), rs[q].converged.begin(), rs[q].converged.end());
          res.stage.insert(res.stage.end(), rs[q].stage.begin(), rs[q].stage.end());
          res.factor_arith = std::max(res.factor_arith, rs[q].factor_arith);
          if (pp) {
            extra_ms += ms[q];
            extra_launches += launches[q];
            extra_flops += profs[q].flops;
            extra_unit_flops += profs[q].unit_flops;
          }
        }
      } else {
        DenseSolveInputs in1 = in;
        in1.graph = s->graphs && !tl_worker_stream ? &s->part_graph[0] : nullptr;
        st = refined ? ldlt_refined_batch(in1, d_omegas.ptr, static_cast<int>(kc), d_z.ptr, res, s->ws,
                                          s->stream, pp)
                     : ldlt_solve_batch(in, d_omegas.ptr, static_cast<int>(kc), d_z.ptr, res, s->ws,
                                        s->stream, pp);
      }
      if (st) return st;
      FH_HIP_TRY(hipMemcpyAsync(z_out + 2 * p * p * c0, d_z.ptr, kc * p * p * 2 * sizeof(double),
                                hipMemcpyDeviceToHost, s->stream));
      FH_HIP_TRY(hipStreamSynchronize(s->stream));
      // the last stage: frequencies no FP32 stage solved, by the plain FP64
      // L D L^T (bordered elimination, no refinement)
      if (refined && (s->refine_stages & FH_REFINE_FP64)) {
        if ((st = fp64_resolve(s, in, omegas.data(), kc, res, z_out + 2 * p * p * c0))) return st;
      }
      factor_arith = std::max(factor_arith, res.factor_arith);
      if (refined)
        for (size_t q = 0; q < kc; ++q) refine_stage = std::max(refine_stage, res.stage[q]);
      for (size_t q = 0; q < kc; ++q) {
        refine_steps_max = std::max(refine_steps_max, res.refine_steps[q]);
        refine_change_max = std::max(refine_change_max, res.refine_change[q]);
        bool bad = res.singular[q] != 0;
        for (size_t e = 0; e < 2 * p * p && !bad; ++e)
          if (!std::isfinite(z_out[2 * (c0 + q) * p * p + e])) bad = true;
        if (bad)
          return set_error(FH_E_SINGULAR,
                           "the mesh system is singular or ill-conditioned at " +
                               std::to_string(freqs[c0 + q]) + " Hz",
                           static_cast<int64_t>(c0 + q), 0, freqs[c0 + q]);
        if (!res.converged[q])
          return set_error(FH_E_NOT_CONVERGED,
                           "iterative refinement did not converge at " +
                               std::to_string(freqs[c0 + q]) + " Hz (relative correction " +
                               std::to_string(res.refine_change[q]) + " after " +
                               std::to_string(res.refine_steps[q]) + " steps)",
                           static_cast<int64_t>(c0 + q), res.refine_steps[q], freqs[c0 + q]);
      }
    }
  }
  if (info) {
    info->nodes = s->counts[0];
    info->segments = s->counts[1];
    info->filaments = s->counts[2];
    info->meshes = s->counts[3];
    info->internal_meshes = s->counts[4];
    info->ports = s->counts[5];
    info->inductance_s = s->inductance_s;
    info->assembly_s = s->assembly_s;
    info->solve_s = seconds_since(t0);
    info->refine_steps_max = refine_steps_max;
    info->factor_arith = factor_arith;
    info->refine_change_max = refine_change_max;
    info->refine_stage = refine_stage;
    info->update_kernel_ms = pp ? prof.total_ms() + extra_ms : 0.0;
    info->update_kernel_launches = pp ? static_cast<int64_t>(prof.launches() + extra_launches) : 0;
    prof.flops += extra_flops;
    prof.unit_flops += extra_unit_flops;
    info->update_kernel_flops = pp ? prof.flops : 0.0;
    info->update_kernel_unit_flops = pp ? prof.unit_flops : 0.0;
    info->solver = FH_SOLVER_DENSE;
    info->gmres_iterations_max = 0;
    info->gmres_iterations = 0;
    info->gmres_residual_max = 0.0;
    info->operator_applies = 0;
    info->operator_apply_ms = 0.0;
  }
  return FH_OK;
}

fh_status fh_system_set_refinement(fh_system *s, uint32_t stages) {
  clear_error();
  if (!s) return set_error(FH_E_INVALID_ARGUMENT, "null system");
  if (stages == 0 || (stages & ~FH_REFINE_ALL))
    return set_error(FH_E_INVALID_ARGUMENT, "stages must be a non-empty mask of FH_REFINE_* bits");
  s->refine_stages = stages;
  return FH_OK;
}

fh_status fh_system_get_schedule(const fh_system *s, fh_schedule *out) {
  clear_error();
  if (!s || !out) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  const bool refined = s->precision == FH_FP32_REFINED;
  const int nb = s->nb_for(refined);
  out->outer_block = nb > 0 ? nb : ldlt_default_nb(refined);
  out->sweep_parts = s->sweep_parts;
  return FH_OK;
}

fh_status fh_system_set_schedule(fh_system *s, const fh_schedule *sched) {
  clear_error();
  if (!s || !sched) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  if (sched->outer_block < 64 || sched->outer_block > 1024 || sched->outer_block % 64 != 0)
    return set_error(FH_E_INVALID_ARGUMENT, "outer_block must be a multiple of 64 from 64 to 1024");
  if (sched->sweep_parts < 1 || sched->sweep_parts > kSweepPartsMax)
    return set_error(FH_E_INVALID_ARGUMENT, "sweep_parts must be 1 to 4");
  if (s->precision == FH_FP32_REFINED)
    s->nb_refined = sched->outer_block;
  else
    s->nb_plain = sched->outer_block;
  s->sweep_parts = sched->sweep_parts;
  s->schedule_fixed = true;
  return FH_OK;
}

fh_status fh_system_autotune(fh_system *s, const double *freqs, size_t K, fh_schedule *chosen,
                             double *best_ms) {
  clear_error();
  if (!s || (!freqs && K)) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  if (s->solver == FH_SOLVER_ITERATIVE || s->basis.internal == 0 || K == 0) {
    if (chosen) (void)fh_system_get_schedule(s, chosen);
    if (best_ms) *best_ms = 0.0;
    return FH_OK;  // nothing to tune: no dense factorization runs
  }
  const bool refined = s->precision == FH_FP32_REFINED;
  const size_t p = s->basis.loops - s->basis.internal;
  std::vector<double> z(2 * K * p * p);
  std::vector<fh_schedule> cand;
  for (const int nb : refined ? std::vector<int>{512, 256, 128} : std::vector<int>{256, 128, 64})
    for (int parts = refined ? 3 : 1; parts >= 1; --parts) cand.push_back(fh_schedule{nb, parts});
  const bool was_fixed = s->schedule_fixed;
  const int nb0 = s->nb_for(refined), parts0 = s->sweep_parts;
  fh_status st;
  // the first candidate needs the most workspace: grow it once, untimed
  s->schedule_fixed = true;
  s->nb_refined = refined ? cand[0].outer_block : s->nb_refined;
  s->nb_plain = refined ? s->nb_plain : cand[0].outer_block;
  s->sweep_parts = cand[0].sweep_parts;
  if ((st = fh_sweep(s, freqs, K, z.data(), nullptr))) {
    s->schedule_fixed = was_fixed;
    return st;
  }
  double best = INFINITY;
  fh_schedule win{nb0 > 0 ? nb0 : ldlt_default_nb(refined), parts0};
  for (const fh_schedule &c : cand) {
    if (refined)
      s->nb_refined = c.outer_block;
    else
      s->nb_plain = c.outer_block;
    s->sweep_parts = c.sweep_parts;
    const auto t0 = Clock::now();
    if ((st = fh_sweep(s, freqs, K, z.data(), nullptr))) {
      s->schedule_fixed = was_fixed;
      return st;
    }
    const double ms = seconds_since(t0) * 1e3;
    if (ms < best) {
      best = ms;
      win = c;
    }
  }
  if (refined)
    s->nb_refined = win.outer_block;
  else
    s->nb_plain = win.outer_block;
  s->sweep_parts = win.sweep_parts;
  s->schedule_fixed = was_fixed;
  tune_store(dense_tune_key(refined, s->basis.internal, K),
             TunedSchedule{win.outer_block, win.sweep_parts, best});
  if (chosen) *chosen = win;
  if (best_ms) *best_ms = best;
  return FH_OK;
}

fh_status fh_system_set_graphs(fh_system *s, int32_t enable) {
  clear_error();
  if (!s) return set_error(FH_E_INVALID_ARGUMENT, "null system");
  s->graphs = enable != 0;
  if (!s->graphs)
    for (auto &g : s->part_graph) g.reset();
  return FH_OK;
}

fh_status fh_system_set_sweep_parts(fh_system *s, int32_t parts) {
  clear_error();
  if (!s) return set_error(FH_E_INVALID_ARGUMENT, "null system");
  if (parts < 1 || parts > kSweepPartsMax) return set_error(FH_E_INVALID_ARGUMENT, "parts must be 1 to 4");
  s->sweep_parts = parts;
  return FH_OK;
}

fh_status fh_impedance(fh_system *s, double f, double *z_out) {
  return fh_sweep(s, &f, 1, z_out, nullptr);
}

fh_status fh_solve(const fh_problem *problem, fh_solver_choice choice, fh_precision precision,
                   const double *freqs, size_t K, double *z_out, fh_sweep_info *info) {
  fh_system *s = nullptr;
  fh_status st = fh_system_create(problem, choice, precision, &s);
  if (st) return st;
  st = fh_sweep(s, freqs, K, z_out, info);
  fh_system_destroy(s);
  return st;
}

fh_status fh_solve_batch(const fh_problem *problems, size_t count, fh_solver_choice choice,
                         fh_precision precision, const double *freqs, size_t K, double *z_out,
                         fh_sweep_info *infos, int32_t concurrency, size_t *failed) {
  clear_error();
  if (failed) *failed = 0;
  if (count == 0) return FH_OK;
  if (!problems || (K && (!freqs || !z_out))) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  for (size_t k = 0; k < K; ++k)
    if (!(std::isfinite(freqs[k]) && freqs[k] >= 0.0))
      return set_error(FH_E_INVALID_FREQUENCY,
                       "frequency " + std::to_string(k) + " is " + std::to_string(freqs[k]) +
                           " Hz; frequencies must be finite and non-negative",
                       static_cast<int64_t>(k), 0, freqs[k]);
  fh_status st = check_device();
  if (st) return st;
  std::vector<size_t> off(count + 1, 0), work(count, 0);
  for (size_t b = 0; b < count; ++b) {
    const size_t p = problems[b].n_ports;
    off[b + 1] = off[b] + 2 * K * p * p;
    for (size_t q = 0; q < problems[b].n_segments && problems[b].grids; ++q)
      work[b] += static_cast<size_t>(problems[b].grids[q].nw) * problems[b].grids[q].nh;
  }
  // longest first: the schedule's tail is a small problem
  std::vector<size_t> order(count);
  std::iota(order.begin(), order.end(), size_t{0});
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return work[a] > work[b]; });
  const size_t want = concurrency > 0 ? static_cast<size_t>(concurrency) : 4;
  const size_t workers = std::min<size_t>(count, want);
  int device = 0;
  FH_HIP_TRY(hipGetDevice(&device));
  struct Outcome {
    fh_status status = FH_OK;
    ErrorState error;
  };
  std::vector<Outcome> outcome(count);
  std::atomic<size_t> next{0};
  auto worker = [&]() {
    (void)hipSetDevice(device);
    hipStream_t stream = nullptr;
    if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) {
      (void)hipGetLastError();
      stream = nullptr;
    }
    tl_worker_stream = stream;
    block_cache(true);
    for (;;) {
      const size_t q = next.fetch_add(1);
      if (q >= count) break;
      const size_t b = order[q];
      fh_system *sys = nullptr;
      fh_status s2 = fh_system_create(&problems[b], choice, precision, &sys);
      if (s2 == FH_OK) {
        s2 = fh_sweep(sys, freqs, K, z_out + off[b], infos ? &infos[b] : nullptr);
        fh_system_destroy(sys);
      }
      outcome[b].status = s2;
      if (s2) outcome[b].error = error_state();
    }
    block_cache(false);
    tl_worker_stream = nullptr;
    if (stream) (void)hipStreamDestroy(stream);
  };
  std::vector<std::thread> pool;
  for (size_t w = 1; w < workers; ++w) {
    try {
      pool.emplace_back(worker);
    } catch (...) {
      break;  // the workers started (this thread at least) take every problem
    }
  }
  worker();
  for (auto &t : pool) t.join();
  for (size_t b = 0; b < count; ++b)
    if (outcome[b].status) {
      error_state() = outcome[b].error;
      error_state().message = "problem " + std::to_string(b) + ": " + error_state().message;
      if (failed) *failed = b;
      return outcome[b].status;
    }
  return FH_OK;
}

fh_status fh_partial_inductance(const fh_filament *filaments, size_t n, double *L_out,
                                int8_t *methods) {
  return fh_partial_inductance_masked(filaments, n, nullptr, L_out, methods);
}

fh_status fh_partial_inductance_masked(const fh_filament *filaments, size_t n, const fh_pair_mask *mask,
                                       double *L_out, int8_t *methods) {
  clear_error();
  if ((n && (!filaments || !L_out))) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  if (mask && mask->n && (!mask->group || (mask->groups && !mask->coupled)))
    return set_error(FH_E_INVALID_ARGUMENT, "null mask array");
  if (n == 0) return FH_OK;
  fh_status st = check_device();
  if (st) return st;
  hipStream_t stream;
  FH_HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{stream};
  DeviceArray<fh_filament> d_f;
  DeviceArray<double> d_L;
  DeviceArray<int8_t> d_m;
  DeviceArray<int32_t> d_g;
  DeviceArray<uint8_t> d_t;
  if ((st = d_f.allocate(n)) || (st = d_L.allocate(n * n))) return st;
  if (methods && (st = d_m.allocate(n * n))) return st;
  FH_HIP_TRY(hipMemcpyAsync(d_f.ptr, filaments, n * sizeof(fh_filament), hipMemcpyHostToDevice, stream));
  DevicePairMask dm;
  if ((st = upload_mask(mask, d_g, d_t, dm, stream))) return st;
  if ((st = fill_inductance(d_f.ptr, n, d_L.ptr, methods ? d_m.ptr : nullptr, stream, mask ? &dm : nullptr)))
    return st;
  FH_HIP_TRY(hipMemcpyAsync(L_out, d_L.ptr, n * n * sizeof(double), hipMemcpyDeviceToHost, stream));
  if (methods)
    FH_HIP_TRY(hipMemcpyAsync(methods, d_m.ptr, n * n, hipMemcpyDeviceToHost, stream));
  FH_HIP_TRY(hipStreamSynchronize(stream));
  return FH_OK;
}

fh_status fh_partial_inductance_noalloc(const fh_filament *d_filaments, size_t n, double *d_L_out,
                                        int8_t *d_methods_out, void *stream) {
  clear_error();
  if (n && (!d_filaments || !d_L_out)) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  fh_status st = check_device();
  if (st) return st;
  return fill_inductance_noalloc(d_filaments, n, d_L_out, d_methods_out,
                                 static_cast<hipStream_t>(stream));
}

fh_status fh_pfft_create(const fh_filament *filaments, size_t n, const fh_pfft_params *params,
                         fh_pfft **out) {
  clear_error();
  if (!params || !out || (n && !filaments)) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  fh_status st = check_device();
  if (st) return st;
  std::unique_ptr<fh_pfft> op(new fh_pfft());
  FH_HIP_TRY(hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking));
  std::vector<fh_filament> f(filaments, filaments + n);
  if ((st = op->op.build(f, *params, op->stream))) return st;
  *out = op.release();
  return FH_OK;
}

fh_status fh_pfft_apply(fh_pfft *op, const double *x, double *y, size_t nvec) {
  clear_error();
  if (!op) return set_error(FH_E_INVALID_ARGUMENT, "null operator");
  const size_t n = op->op.size();
  if (n == 0 || nvec == 0) return FH_OK;
  if (!x || !y) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  if (nvec > 1024) return set_error(FH_E_INVALID_ARGUMENT, "at most 1024 vectors per call");
  fh_status st;
  if (op->dx.count < n * nvec && ((st = op->dx.allocate(n * nvec)) || (st = op->dy.allocate(n * nvec))))
    return st;
  FH_HIP_TRY(hipMemcpyAsync(op->dx.ptr, x, n * nvec * sizeof(double), hipMemcpyHostToDevice, op->stream));
  if ((st = op->op.apply(op->dx.ptr, op->dy.ptr, static_cast<int>(nvec), op->stream))) return st;
  FH_HIP_TRY(hipMemcpyAsync(y, op->dy.ptr, n * nvec * sizeof(double), hipMemcpyDeviceToHost, op->stream));
  FH_HIP_TRY(hipStreamSynchronize(op->stream));
  return FH_OK;
}

fh_status fh_pfft_get_stats(const fh_pfft *op, fh_pfft_stats *out) {
  clear_error();
  if (!op || !out) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  *out = op->op.stats();
  return FH_OK;
}

fh_status fh_pfft_data(const fh_pfft *op, int64_t *proj_off, uint32_t *proj_idx, double *proj_w,
                       int64_t *near_off, uint32_t *near_col, double *near_val) {
  clear_error();
  if (!op) return set_error(FH_E_INVALID_ARGUMENT, "null operator");
  return op->op.copy_data(proj_off, proj_idx, proj_w, near_off, near_col, near_val);
}

void fh_pfft_destroy(fh_pfft *op) { delete op; }

fh_status fh_gmres_dense(size_t n, const double *a, const double *b, double *x,
                         const fh_gmres_params *params, int32_t complex_arith,
                         uint64_t *iterations, double *relative_residual, int32_t *converged) {
  clear_error();
  if (!params || (n && (!a || !b || !x))) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  fh_status st = check_device();
  if (st) return st;
  GmresOutcome o;
  if ((st = gmres_dense(n, a, b, x, *params, complex_arith != 0, o))) return st;
  if (iterations) *iterations = o.iterations;
  if (relative_residual) *relative_residual = o.relative_residual;
  if (converged) *converged = o.converged ? 1 : 0;
  return FH_OK;
}

fh_status fh_device_math(const double *x, size_t n, double *log_out, double *atan_out) {
  clear_error();
  if (n && (!x || !log_out || !atan_out)) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  fh_status st = check_device();
  if (st) return st;
  return device_math(x, n, log_out, atan_out);
}

// Gradients (no counterpart in the reference): Z_ab = i_a^T Z_f i_b with
// Z_f = diag(R) + jw L over filaments and i_a = M^T lambda_a, lambda_a =
// [-x_a; e_a] the loop currents of port a (x = Zee^-1 Zep).  Z is stationary
// in lambda, so dZ_ab/dp = i_a^T (dR/dp + jw dL/dp) i_b.
fh_status fh_grad(fh_system *s, const fh_param_map *map, const double *freqs, size_t K, double *z_out,
                  double *dR_dp, double *dL_dp) {
  clear_error();
  if (!s || !map || (K && !freqs)) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  const size_t P = map->n_params;
  if (K && P && (!dR_dp || !dL_dp)) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  if (s->iter || s->solver != FH_SOLVER_DENSE)
    return set_error(FH_E_UNSUPPORTED, "gradients need a dense system");
  for (size_t k = 0; k < K; ++k)
    if (!(std::isfinite(freqs[k]) && freqs[k] >= 0.0))
      return set_error(FH_E_INVALID_FREQUENCY,
                       "frequency " + std::to_string(k) + " is " + std::to_string(freqs[k]) +
                           " Hz; frequencies must be finite and non-negative",
                       static_cast<int64_t>(k), 0, freqs[k]);
  const size_t n = s->filaments.size(), m = s->basis.internal, p = s->basis.loops - m;
  if (K == 0) return FH_OK;
  if (p == 0) return set_error(FH_E_INVALID_ARGUMENT, "the problem has no ports");
  const size_t per_pass = 128 / (p * p);
  if (per_pass == 0) return set_error(FH_E_UNSUPPORTED, "gradients support at most 11 ports");
  const fh_problem prob = s->problem();
  // Filament tangents along every direction (host), then dR_f/dp.
  std::vector<fh_filament> T(P * n);
  for (size_t q = 0; q < P; ++q) {
    std::vector<fh_filament> t;
    fh_status st = discretize_tangent(
        prob, map->d_nodes ? map->d_nodes + q * 3 * prob.n_nodes : nullptr,
        map->d_widths ? map->d_widths + q * prob.n_segments : nullptr,
        map->d_heights ? map->d_heights + q * prob.n_segments : nullptr, t);
    if (st) return st;
    if (t.size() != n) return set_error(FH_E_INVALID_ARGUMENT, "tangent size mismatch");
    std::copy(t.begin(), t.end(), T.begin() + q * n);
  }
  // Loop currents: x = Zee^-1 Zep of every frequency, refined to working
  // accuracy (they enter the gradient to first order).
  std::vector<double> X;
  {
    fh_status st = refined_x(s, freqs, K, X);
    if (st) return st;
  }
  // Filament currents I[k][a][f] = sum_l M[l][f] lambda_a[l].
  std::vector<double2> I(K * p * n, make_double2(0.0, 0.0));
  const LoopBasis &B = s->basis;
  for (size_t k = 0; k < K; ++k)
    for (size_t a = 0; a < p; ++a) {
      double2 *Ia = I.data() + (k * p + a) * n;
      const double *xr = X.data() + (2 * (k * p + a)) * m, *xi = xr + m;
      for (size_t l = 0; l < B.loops; ++l) {
        double lr, li;
        if (l < m) {
          lr = -xr[l];
          li = -xi[l];
        } else {
          lr = l - m == a ? 1.0 : 0.0;
          li = 0.0;
        }
        if (lr == 0.0 && li == 0.0) continue;
        for (int64_t e = B.row_ptr[l]; e < B.row_ptr[l + 1]; ++e) {
          double2 &c = Ia[B.branch[e]];
          c.x += B.sign[e] * lr;
          c.y += B.sign[e] * li;
        }
      }
    }
  DeviceArray<double2> d_I;
  fh_status st;
  if ((st = d_I.allocate(K * p * n))) return st;
  FH_HIP_TRY(hipMemcpyAsync(d_I.ptr, I.data(), K * p * n * sizeof(double2), hipMemcpyHostToDevice, s->stream));
  // Z in the stationary form Z_ab = i_a^T (R + jw L) i_b: the same value as
  // the Schur complement for exact currents, without its cancellation
  // (Zpp and Zep^T x can exceed Z by orders of magnitude at low frequency),
  // so Z(p) is smooth to rounding and finite differences of it are clean.
  if (z_out) {
    std::vector<double2> LI;
    if ((st = inductance_products(s->d_L.ptr, static_cast<int>(n), d_I.ptr, static_cast<int>(K * p), LI,
                                  s->stream)))
      return st;
    for (size_t k = 0; k < K; ++k) {
      const double omega = 6.283185307179586 * freqs[k];
      for (size_t a = 0; a < p; ++a)
        for (size_t b = 0; b < p; ++b) {
          const double2 *Ia = I.data() + (k * p + a) * n, *Ib = I.data() + (k * p + b) * n;
          const double2 *Lb = LI.data() + (k * p + b) * n;
          double rr = 0.0, ri = 0.0, lr = 0.0, li = 0.0;
          for (size_t f = 0; f < n; ++f) {
            const double R = s->resistances[f];
            rr += R * (Ia[f].x * Ib[f].x - Ia[f].y * Ib[f].y);
            ri += R * (Ia[f].x * Ib[f].y + Ia[f].y * Ib[f].x);
            lr += Ia[f].x * Lb[f].x - Ia[f].y * Lb[f].y;
            li += Ia[f].x * Lb[f].y + Ia[f].y * Lb[f].x;
          }
          // R part + jw (lr + j li)
          z_out[2 * ((k * p + a) * p + b)] = rr - omega * li;
          z_out[2 * ((k * p + a) * p + b) + 1] = freqs[k] == 0.0 ? 0.0 : ri + omega * lr;
        }
    }
  }
  if (P == 0) return FH_OK;
  // dL contracted with the currents on the device, frequencies in passes of
  // at most 128 / p^2.
  DeviceArray<fh_filament> d_T;
  if ((st = d_T.allocate(P * n))) return st;
  FH_HIP_TRY(hipMemcpyAsync(d_T.ptr, T.data(), P * n * sizeof(fh_filament), hipMemcpyHostToDevice, s->stream));
  std::vector<double2> dLc(K * P * p * p);  // [k][q][a][b]
  for (size_t k0 = 0; k0 < K; k0 += per_pass) {
    const size_t kc = std::min(per_pass, K - k0);
    std::vector<double2> c;
    if ((st = contract_inductance_derivative(s->d_filaments.ptr, static_cast<int>(n), d_T.ptr, static_cast<int>(P),
                                             d_I.ptr + k0 * p * n, static_cast<int>(kc), static_cast<int>(p), c,
                                             s->stream)))
      return st;
    for (size_t q = 0; q < P; ++q)
      for (size_t k = 0; k < kc; ++k)
        for (size_t ab = 0; ab < p * p; ++ab)
          dLc[((k0 + k) * P + q) * p * p + ab] = c[(q * kc + k) * p * p + ab];
  }
  // dZ = sum_f I_a I_b dR_f + jw (contraction); dR = Re dZ, dL = Im dZ / w.
  for (size_t q = 0; q < P; ++q) {
    std::vector<double> dr(n);
    for (size_t f = 0; f < n; ++f) {
      const fh_filament &F = s->filaments[f], &t = T[q * n + f];
      dr[f] = s->resistances[f] * (t.length / F.length - t.width / F.width - t.height / F.height);
    }
    for (size_t k = 0; k < K; ++k) {
      const double omega = 6.283185307179586 * freqs[k];
      for (size_t a = 0; a < p; ++a)
        for (size_t b = 0; b < p; ++b) {
          const double2 *Ia = I.data() + (k * p + a) * n, *Ib = I.data() + (k * p + b) * n;
          double re = 0.0, im = 0.0;
          for (size_t f = 0; f < n; ++f) {
            if (dr[f] == 0.0) continue;
            re += dr[f] * (Ia[f].x * Ib[f].x - Ia[f].y * Ib[f].y);
            im += dr[f] * (Ia[f].x * Ib[f].y + Ia[f].y * Ib[f].x);
          }
          const double2 c = dLc[(k * P + q) * p * p + a * p + b];
          // + jw (c.x + j c.y)
          re -= omega * c.y;
          im += omega * c.x;
          const size_t o = ((k * P + q) * p + a) * p + b;
          dR_dp[o] = re;
          dL_dp[o] = freqs[k] == 0.0 ? 0.0 : im / omega;
        }
    }
  }
  return FH_OK;
}

fh_status fh_partial_inductance_derivative(const fh_filament *filaments, size_t n, const fh_filament *tangents,
                                           size_t n_dirs, const uint32_t *pairs, size_t count, double *value_out,
                                           double *deriv_out) {
  clear_error();
  if (count == 0) return FH_OK;
  if (!filaments || !tangents || !pairs || !value_out || !deriv_out)
    return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  if (n >= (size_t(1) << 30)) return set_error(FH_E_INVALID_ARGUMENT, "too many filaments");
  std::vector<uint4> pr(count);
  for (size_t k = 0; k < count; ++k) {
    const uint32_t f = pairs[3 * k], g = pairs[3 * k + 1], q = pairs[3 * k + 2];
    if (f >= n || g >= n || q >= n_dirs)
      return set_error(FH_E_INVALID_ARGUMENT, "pair " + std::to_string(k) + " is out of range",
                       static_cast<int64_t>(k));
    pr[k] = make_uint4(f, g, q, 0u);
  }
  fh_status st = check_device();
  if (st) return st;
  hipStream_t stream = nullptr;
  FH_HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{stream};
  DeviceArray<fh_filament> d_F, d_T;
  DeviceArray<uint4> d_p;
  DeviceArray<double> d_v, d_d;
  if ((st = d_F.allocate(n)) || (st = d_T.allocate(n * n_dirs)) || (st = d_p.allocate(count)) ||
      (st = d_v.allocate(count)) || (st = d_d.allocate(count)))
    return st;
  FH_HIP_TRY(hipMemcpyAsync(d_F.ptr, filaments, n * sizeof(fh_filament), hipMemcpyHostToDevice, stream));
  FH_HIP_TRY(hipMemcpyAsync(d_T.ptr, tangents, n * n_dirs * sizeof(fh_filament), hipMemcpyHostToDevice, stream));
  FH_HIP_TRY(hipMemcpyAsync(d_p.ptr, pr.data(), count * sizeof(uint4), hipMemcpyHostToDevice, stream));
  if ((st = pair_derivatives(d_F.ptr, static_cast<int>(n), d_T.ptr, d_p.ptr, count, d_v.ptr, d_d.ptr, stream)))
    return st;
  FH_HIP_TRY(hipMemcpyAsync(value_out, d_v.ptr, count * sizeof(double), hipMemcpyDeviceToHost, stream));
  FH_HIP_TRY(hipMemcpyAsync(deriv_out, d_d.ptr, count * sizeof(double), hipMemcpyDeviceToHost, stream));
  FH_HIP_TRY(hipStreamSynchronize(stream));
  for (size_t k = 0; k < count; ++k) {
    if (std::isnan(value_out[k]))
      return set_error(FH_E_KERNEL, "a sampled-filament pair has coincident sample lines", pairs[3 * k],
                       pairs[3 * k + 1], 0.0, 1);
    if (std::isnan(deriv_out[k]))
      return set_error(FH_E_UNSUPPORTED,
                       "pair " + std::to_string(k) +
                           ": the direction turns a perpendicular or parallel filament pair relative to each other",
                       pairs[3 * k], pairs[3 * k + 1]);
  }
  return FH_OK;
}

fh_status fh_currents(fh_system *s, double frequency, const double *excitation, size_t n_excitation,
                      double *filament_current, double *current_density, double *power_w,
                      double *segment_current, double *segment_power_w, int32_t *conductor_of_segment,
                      size_t *n_conductors, int64_t *peak_filament, double *port_voltage,
                      fh_power_balance *balance) {
  clear_error();
  if (!s) return set_error(FH_E_INVALID_ARGUMENT, "null system");
  if (!(std::isfinite(frequency) && frequency >= 0.0))  // check_frequency (solve.rs:992+)
    return set_error(FH_E_INVALID_FREQUENCY,
                     "frequency 0 is " + std::to_string(frequency) +
                         " Hz; frequencies must be finite and non-negative",
                     0, 0, frequency);
  const size_t n = s->filaments.size(), m = s->basis.internal, p = s->basis.loops - m;
  bool ok = n_excitation == p && (p == 0 || excitation);  // check_excitation (solve.rs:976-990)
  for (size_t e = 0; ok && e < 2 * p; ++e) ok = std::isfinite(excitation[e]);
  if (!ok)
    return set_error(FH_E_INVALID_EXCITATION,
                     "excitation must have one finite entry per port (" + std::to_string(p) + "), got " +
                         std::to_string(n_excitation) + " entries",
                     static_cast<int64_t>(p), static_cast<int64_t>(n_excitation));
  using cd = std::complex<double>;
  auto cmul = [](cd a, cd b) {
    return cd(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
  };
  const double omega = 6.283185307179586 * frequency;
  std::vector<cd> ex(p), internal(m), voltage(p);
  for (size_t a = 0; a < p; ++a) ex[a] = cd(excitation[2 * a], excitation[2 * a + 1]);
  if (s->iter) {
    std::vector<double> in, v;
    GmresOutcome o;
    fh_status st = s->iter->currents(frequency, excitation, in, v, o);
    if (st) return st;
    for (size_t i = 0; i < m; ++i) internal[i] = cd(in[2 * i], in[2 * i + 1]);
    for (size_t a = 0; a < p; ++a) voltage[a] = cd(v[2 * a], v[2 * a + 1]);
  } else {
    std::vector<double> X;
    fh_status st = refined_x(s, &frequency, 1, X);
    if (st) return st;
    // internal = -x I_p
    for (size_t i = 0; i < m; ++i) {
      cd acc(0.0, 0.0);
      for (size_t a = 0; a < p; ++a)
        acc += cmul(cd(X[(2 * a) * m + i], X[(2 * a + 1) * m + i]), ex[a]);
      internal[i] = -acc;
    }
    // V_p = Zpp I_p + Zep^T internal, from the port rows of the mesh blocks
    const size_t loops = s->basis.loops;
    std::vector<double> lrow(p * loops), rrow(p * loops, 0.0);
    if (p)
      FH_HIP_TRY(hipMemcpy(lrow.data(), s->d_lmesh.ptr + m * loops, lrow.size() * sizeof(double),
                           hipMemcpyDeviceToHost));
    for (size_t e = 0; e < s->r_vals.size(); ++e)
      if (static_cast<size_t>(s->r_rows[e]) >= m)
        rrow[(static_cast<size_t>(s->r_rows[e]) - m) * loops + s->r_cols[e]] = s->r_vals[e];
    const bool dc = frequency == 0.0;
    auto z = [&](size_t a, size_t j) {  // Z[m + a][j], symmetric
      return cd(rrow[a * loops + j], dc ? 0.0 : omega * lrow[a * loops + j]);
    };
    for (size_t a = 0; a < p; ++a) {
      cd acc(0.0, 0.0);
      for (size_t b = 0; b < p; ++b) acc += cmul(z(a, m + b), ex[b]);
      cd tr(0.0, 0.0);
      for (size_t i = 0; i < m; ++i) tr += cmul(z(a, i), internal[i]);
      voltage[a] = acc + tr;
    }
  }
  for (size_t i = 0; i < m; ++i)
    if (!(std::isfinite(internal[i].real()) && std::isfinite(internal[i].imag())))
      return set_error(FH_E_SINGULAR,
                       "the mesh system is singular or ill-conditioned at " + std::to_string(frequency) + " Hz",
                       0, 0, frequency);
  for (size_t a = 0; a < p; ++a)
    if (!(std::isfinite(voltage[a].real()) && std::isfinite(voltage[a].imag())))
      return set_error(FH_E_SINGULAR,
                       "the mesh system is singular or ill-conditioned at " + std::to_string(frequency) + " Hz",
                       0, 0, frequency);
  // branch = M^T [internal; excitation] (currents.rs:200-206)
  std::vector<cd> branch(n, cd(0.0, 0.0));
  const LoopBasis &B = s->basis;
  for (size_t l = 0; l < B.loops; ++l) {
    const cd x = l < m ? internal[l] : ex[l - m];
    for (int64_t e = B.row_ptr[l]; e < B.row_ptr[l + 1]; ++e) branch[B.branch[e]] += x * B.sign[e];
  }
  // Layout (currents.rs:144-184): conductors by union of segment end nodes
  const size_t nodes = s->prob_nodes.size() / 3, nseg = s->prob_segments.size();
  std::vector<size_t> parent(nodes);
  std::iota(parent.begin(), parent.end(), size_t(0));
  auto find = [&](size_t x) {
    while (parent[x] != x) {
      parent[x] = parent[parent[x]];
      x = parent[x];
    }
    return x;
  };
  for (const fh_segment &d : s->prob_segments) {
    const size_t a = find(d.a), b = find(d.b);
    if (a != b) parent[a] = b;
  }
  std::vector<size_t> roots;
  std::vector<int32_t> conductor(nseg);
  for (size_t q = 0; q < nseg; ++q) {
    const size_t r = find(s->prob_segments[q].a);
    size_t c = std::find(roots.begin(), roots.end(), r) - roots.begin();
    if (c == roots.size()) roots.push_back(r);
    conductor[q] = static_cast<int32_t>(c);
  }
  std::vector<cd> seg_current(nseg, cd(0.0, 0.0));
  std::vector<double> seg_power(nseg, 0.0), peak(roots.size(), -1.0);
  std::vector<int64_t> peak_at(roots.size(), -1);
  double dissipated = 0.0;
  size_t f = 0;
  for (size_t q = 0; q < nseg; ++q) {
    const size_t per = static_cast<size_t>(s->prob_grids[q].nw) * s->prob_grids[q].nh;
    for (size_t k = 0; k < per; ++k, ++f) {
      const fh_filament &F = s->filaments[f];
      const cd i = branch[f];
      const double density = std::hypot(i.real(), i.imag()) / (F.width * F.height);
      const double power = 0.5 * (i.real() * i.real() + i.imag() * i.imag()) * s->resistances[f];
      dissipated += power;
      seg_current[q] += i;
      seg_power[q] += power;
      const int32_t c = conductor[q];
      if (peak_at[c] < 0 || density > peak[c]) {
        peak[c] = density;
        peak_at[c] = static_cast<int64_t>(f);
      }
      if (filament_current) {
        filament_current[2 * f] = i.real();
        filament_current[2 * f + 1] = i.imag();
      }
      if (current_density) current_density[f] = density;
      if (power_w) power_w[f] = power;
    }
  }
  double input = 0.0;
  for (size_t a = 0; a < p; ++a) input += (std::conj(ex[a]) * voltage[a]).real();
  input *= 0.5;
  const double scale = std::max(dissipated, input);
  for (size_t q = 0; q < nseg; ++q) {
    if (segment_current) {
      segment_current[2 * q] = seg_current[q].real();
      segment_current[2 * q + 1] = seg_current[q].imag();
    }
    if (segment_power_w) segment_power_w[q] = seg_power[q];
    if (conductor_of_segment) conductor_of_segment[q] = conductor[q];
  }
  if (n_conductors) *n_conductors = roots.size();
  if (peak_filament)
    for (size_t c = 0; c < roots.size(); ++c) peak_filament[c] = peak_at[c];
  for (size_t a = 0; a < p && port_voltage; ++a) {
    port_voltage[2 * a] = voltage[a].real();
    port_voltage[2 * a + 1] = voltage[a].imag();
  }
  if (balance) {
    balance->dissipated_w = dissipated;
    balance->input_w = input;
    balance->relative_residual = scale > 0.0 ? std::fabs(dissipated - input) / scale : 0.0;
  }
  return FH_OK;
}

}  // extern "C"
