// Matrix-free mesh system on the GPU (fh_iterative.hip): restarted GMRES on
// device vectors and the pFFT-based impedance of iterative.rs.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <vector>

#include "fh_internal.hpp"
#include "fh_pfft.hpp"

namespace fh {

// GmresOutcome (gmres.rs:63-73).
struct GmresOutcome {
  size_t iterations = 0;
  double relative_residual = 0.0;
  bool converged = true;
};

// gmres (gmres.rs:91-206): restarted, right-preconditioned by a diagonal
// (Jacobi) or the identity, modified Gram-Schmidt applied twice, complex
// Givens rotations.  Vectors live in device memory; the (restart + 1) x
// restart Hessenberg system and its rotations live on the host, one
// copy of the new column per iteration.  Reductions are deterministic.
class DeviceGmres {
 public:
  using Op = std::function<fh_status(const double2 *in, double2 *out)>;
  // b, x: n complex values (device); x holds the initial guess.  d_diag
  // (device, n) or null for no preconditioner.  complex_arith = false is
  // the real form (every imaginary part zero, norms of real parts).
  fh_status solve(const Op &apply, const double2 *d_diag, const double2 *d_b, double2 *d_x,
                  size_t n, const fh_gmres_params &params, bool complex_arith, hipStream_t stream,
                  GmresOutcome &out);
  size_t bytes() const;

 private:
  fh_status reserve(size_t n, size_t restart);
  fh_status norm(const double2 *d_v, size_t n, bool complex_arith, int slot, hipStream_t stream);
  size_t n_ = 0, restart_ = 0;
  DeviceArray<double2> basis_, w_, tmp_, residual_, col_, y_;
  DeviceArray<double2> part_a_, part_b_;
  DeviceArray<double> mx_, ss_, scalars_;
};

// IterativeSystem (iterative.rs:155-536): port columns through the
// operator, Jacobi diagonals from exact loop inductances, Z by one GMRES per
// port.
class IterativeSystem {
 public:
  fh_status build(const std::vector<fh_filament> &filaments, const std::vector<double> &resistances,
                  const LoopBasis &basis, const fh_iterative_params &params, hipStream_t stream);
  // z: p x p complex {re, im} row-major (host).  iterations: p entries.
  fh_status impedance(double frequency, double *z, std::vector<size_t> &iterations,
                      double &residual_max);
  // solve_currents (solve.rs:1110-1149) with GMRES (iterative.rs:329-380):
  // excitation p complex {re, im}; internal (m) = -Zee^-1 Zep I_p and the
  // port voltages (p) Zpp I_p + Zep^T internal, complex {re, im}.
  fh_status currents(double frequency, const double *excitation, std::vector<double> &internal,
                     std::vector<double> &voltage, GmresOutcome &outcome);
  const PfftOperator &op() const { return op_; }
  PfftOperator &op() { return op_; }
  double operator_seconds() const { return operator_s_; }
  // applies of the operator (each on V = 2 vectors for an AC product) and
  // their total time when profiling (HIP events), for measurement
  void set_profiling(bool on) { profiling_ = on; }
  double apply_ms() const { return apply_ms_; }
  size_t applies() const { return applies_; }
  void reset_counters() {
    apply_ms_ = 0.0;
    applies_ = 0;
  }
  size_t bytes() const;

 private:
  fh_status apply_ee(double omega, bool dc, const double2 *x, double2 *y);
  hipStream_t stream_ = nullptr;
  size_t n_ = 0, m_ = 0, p_ = 0, loops_ = 0;
  fh_gmres_params gmres_{};
  PfftOperator op_;
  DeviceGmres solver_;
  std::vector<double> r_ep_, l_ep_, r_pp_, l_pp_, r_diag_, l_diag_;  // row-major
  DeviceArray<double> d_res_;
  DeviceArray<int64_t> d_row_ptr_, d_tr_ptr_;
  DeviceArray<int32_t> d_branch_, d_tr_loop_;
  DeviceArray<double> d_sign_, d_tr_sign_;
  DeviceArray<double> d_cur_, d_flux_;  // 2 x n branch vectors
  DeviceArray<double2> d_diag_, d_rhs_, d_x_;
  double operator_s_ = 0.0;
  bool profiling_ = false;
  double apply_ms_ = 0.0;
  size_t applies_ = 0;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;

 public:
  ~IterativeSystem();
};

// GMRES on a dense column-major complex matrix (device copy of a host
// matrix), no preconditioner.
fh_status gmres_dense(size_t n, const double *a, const double *b, double *x,
                      const fh_gmres_params &params, bool complex_arith, GmresOutcome &out);

}  // namespace fh
