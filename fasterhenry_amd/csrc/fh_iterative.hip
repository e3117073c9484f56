// Restarted GMRES on device vectors (gmres.rs:91-257) and the matrix-free
// mesh system of iterative.rs:155-536 on the pFFT operator.
//
// GMRES: the Krylov basis, the work vectors and the preconditioner live in
// device memory; an iteration is one operator product, 2 (j + 1) fused
// Gram-Schmidt launches (each applies the previous projection and computes
// the next inner product over the same elements), a scaled norm, and one
// device-to-host copy of the new Hessenberg column.  Inner products and
// norms reduce per block in a fixed tree and then over the blocks in a
// fixed order, so the iteration is a deterministic function of its inputs.
//
// Mesh system: Z_ee x = M_e (R + jwL) M_e^T x with L the pFFT operator.
// M_e^T (scatter) runs one thread per branch over the loop basis
// transposed (loops ascending, the order iterative.rs:405-416 adds them);
// M_e (gather) one thread per loop over its row, with the branch voltage
// drop r i - w flux formed on the fly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>

#include "fh_iterative.hpp"

namespace fh {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 256;
constexpr size_t kExactDiagonalMaxLoop = 32;  // iterative.rs:94

int reduction_blocks(size_t n) {
  return static_cast<int>(std::min<size_t>(kMaxBlocks, std::max<size_t>(1, (n + 2047) / 2048)));
}

__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// num-complex division: (a conj(b)) / |b|^2 with |b|^2 = re^2 + im^2
__device__ inline double2 cdiv(double2 a, double2 b) {
  const double n = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / n, (a.y * b.x - a.x * b.y) / n);
}

// Fixed-tree block reductions over kThreads threads (4 wavefronts).
__device__ inline double2 block_sum(double2 v, double2 *lds) {
  for (int d = 32; d > 0; d >>= 1) {
    v.x += __shfl_xor(v.x, d);
    v.y += __shfl_xor(v.y, d);
  }
  const int wave = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) lds[wave] = v;
  __syncthreads();
  return cadd(cadd(lds[0], lds[1]), cadd(lds[2], lds[3]));
}

__device__ inline double block_max(double v, double *lds) {
  for (int d = 32; d > 0; d >>= 1) v = fmax(v, __shfl_xor(v, d));
  const int wave = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) lds[wave] = v;
  __syncthreads();
  return fmax(fmax(lds[0], lds[1]), fmax(lds[2], lds[3]));
}

__device__ inline double block_sum_real(double v, double *lds) {
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d);
  const int wave = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) lds[wave] = v;
  __syncthreads();
  return (lds[0] + lds[1]) + (lds[2] + lds[3]);
}

// Sum of `count` (<= kThreads) block partials, the same in every block.
__device__ inline double2 reduce_partials(const double2 *part, int count, double2 *lds) {
  const double2 v = threadIdx.x < count ? part[threadIdx.x] : make_double2(0.0, 0.0);
  return block_sum(v, lds);
}

// One modified Gram-Schmidt step: with `v_prev`, w -= h v_prev where h is
// the sum of `part_prev` (block 0 adds h to col[col_idx]); with `v_next`,
// the block's partial of conj(v_next) . w into part_out; without it, the
// block's partial max |w| into mx_out (the first pass of the norm).
__global__ void __launch_bounds__(kThreads) mgs_kernel(
    double2 *__restrict__ w, const double2 *__restrict__ v_prev,
    const double2 *__restrict__ part_prev, int nparts, double2 *__restrict__ col, int col_idx,
    const double2 *__restrict__ v_next, double2 *__restrict__ part_out, double *__restrict__ mx_out,
    size_t n, int complex_arith) {
  __shared__ double2 lds[4];
  __shared__ double ldsr[4];
  double2 h = make_double2(0.0, 0.0);
  if (v_prev) {
    h = reduce_partials(part_prev, nparts, lds);
    if (blockIdx.x == 0 && threadIdx.x == 0) col[col_idx] = cadd(col[col_idx], h);
  }
  const double2 neg = make_double2(-h.x, -h.y);
  double2 acc = make_double2(0.0, 0.0);
  double mx = 0.0;
  const size_t stride = static_cast<size_t>(gridDim.x) * kThreads;
  for (size_t i = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x; i < n; i += stride) {
    double2 wi = w[i];
    if (v_prev) {
      wi = cadd(wi, cmul(neg, v_prev[i]));
      w[i] = wi;
    }
    if (v_next) {
      acc = cadd(acc, cmul(cconj(v_next[i]), wi));
    } else {
      mx = fmax(mx, complex_arith ? hypot(wi.x, wi.y) : fabs(wi.x));
    }
  }
  if (v_next) {
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) part_out[blockIdx.x] = acc;
  } else {
    mx = block_max(mx, ldsr);
    if (threadIdx.x == 0) mx_out[blockIdx.x] = mx;
  }
}

__global__ void __launch_bounds__(kThreads) max_kernel(const double2 *__restrict__ v, size_t n,
                                                        int complex_arith, double *__restrict__ mx_out) {
  __shared__ double ldsr[4];
  double mx = 0.0;
  const size_t stride = static_cast<size_t>(gridDim.x) * kThreads;
  for (size_t i = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x; i < n; i += stride) {
    const double2 vi = v[i];
    mx = fmax(mx, complex_arith ? hypot(vi.x, vi.y) : fabs(vi.x));
  }
  mx = block_max(mx, ldsr);
  if (threadIdx.x == 0) mx_out[blockIdx.x] = mx;
}

// gmres.rs:209-226 second pass: sum of |v / scale|^2.
__global__ void __launch_bounds__(kThreads) sumsq_kernel(const double2 *__restrict__ v, size_t n,
                                                          int complex_arith,
                                                          const double *__restrict__ mx, int nparts,
                                                          double *__restrict__ ss_out) {
  __shared__ double ldsr[4];
  const double scale = block_max(threadIdx.x < nparts ? mx[threadIdx.x] : 0.0, ldsr);
  double s = 0.0;
  if (scale != 0.0 && isfinite(scale)) {
    const size_t stride = static_cast<size_t>(gridDim.x) * kThreads;
    for (size_t i = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x; i < n; i += stride) {
      const double2 u = make_double2(v[i].x / scale, v[i].y / scale);
      s += complex_arith ? u.x * u.x + u.y * u.y : u.x * u.x;
    }
  }
  s = block_sum_real(s, ldsr);
  if (threadIdx.x == 0) ss_out[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kThreads) norm_finish_kernel(const double *__restrict__ mx,
                                                                const double *__restrict__ ss,
                                                                int nparts, double *__restrict__ out) {
  __shared__ double ldsr[4];
  const double scale = block_max(threadIdx.x < nparts ? mx[threadIdx.x] : 0.0, ldsr);
  const double sum = block_sum_real(threadIdx.x < nparts ? ss[threadIdx.x] : 0.0, ldsr);
  if (threadIdx.x == 0) out[0] = (scale == 0.0 || !isfinite(scale)) ? scale : scale * sqrt(sum);
}

__global__ void scale_kernel(const double2 *__restrict__ v, const double *__restrict__ s,
                             double2 *__restrict__ out, size_t n) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i < n) {
    const double d = s[0];
    out[i] = make_double2(v[i].x / d, v[i].y / d);
  }
}

__global__ void precondition_kernel(const double2 *__restrict__ v, const double2 *__restrict__ d,
                                    double2 *__restrict__ out, size_t n) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i < n) out[i] = cdiv(v[i], d[i]);
}

__global__ void residual_kernel(const double2 *__restrict__ b, double2 *__restrict__ r, size_t n) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i < n) r[i] = csub(b[i], r[i]);
}

// x += P^-1 (sum_k y_k basis_k), the update in basis order from zero.
__global__ void update_kernel(const double2 *__restrict__ basis, const double2 *__restrict__ y,
                              int steps, const double2 *__restrict__ d, double2 *__restrict__ x,
                              size_t n) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  double2 u = make_double2(0.0, 0.0);
  for (int k = 0; k < steps; ++k) u = cadd(u, cmul(y[k], basis[static_cast<size_t>(k) * n + i]));
  if (d) u = cdiv(u, d[i]);
  x[i] = cadd(x[i], u);
}

__global__ void zero2_kernel(double2 *__restrict__ p, size_t n) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i < n) p[i] = make_double2(0.0, 0.0);
}

// Dense column-major complex product (the oracle's dense_apply order).
__global__ void dense_apply_kernel(const double2 *__restrict__ a, const double2 *__restrict__ x,
                                   double2 *__restrict__ y, size_t n) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  double2 s = make_double2(0.0, 0.0);
  for (size_t k = 0; k < n; ++k) s = cadd(s, cmul(a[k * n + i], x[k]));
  y[i] = s;
}

// Branch currents M_e^T x: cur[b] (re) and cur[n + b] (im).
__global__ void scatter_kernel(const int64_t *__restrict__ tr_ptr, const int32_t *__restrict__ tr_loop,
                               const double *__restrict__ tr_sign, const double2 *__restrict__ x,
                               size_t n, double *__restrict__ cur) {
  const size_t b = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (b >= n) return;
  double re = 0.0, im = 0.0;
  for (int64_t e = tr_ptr[b]; e < tr_ptr[b + 1]; ++e) {
    const double2 xi = x[tr_loop[e]];
    const double s = tr_sign[e];
    re += xi.x * s;
    im += xi.y * s;
  }
  cur[b] = re;
  cur[n + b] = im;
}

// y_i = sum over row i of drop(b) sign with drop = r i - w flux_im + j (r i_im + w flux_re)
// (iterative.rs:432-448) or, at DC, r i (iterative.rs:423-429).
__global__ void gather_kernel(const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ branch,
                              const double *__restrict__ sign, const double *__restrict__ res,
                              const double *__restrict__ cur, const double *__restrict__ flux,
                              double omega, int dc, size_t n, size_t m, double2 *__restrict__ y) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i >= m) return;
  double2 sum = make_double2(0.0, 0.0);
  for (int64_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
    const int32_t b = branch[e];
    const double r = res[b];
    double dre, dim;
    if (dc) {
      dre = cur[b] * r;
      dim = cur[n + b] * r;
    } else {
      dre = r * cur[b] - omega * flux[n + b];
      dim = r * cur[n + b] + omega * flux[b];
    }
    const double s = sign[e];
    sum = make_double2(sum.x + dre * s, sum.y + dim * s);
  }
  y[i] = sum;
}

unsigned grid_for(size_t n) { return static_cast<unsigned>((n + kThreads - 1) / kThreads); }

// gmres.rs:246-257
void givens(std::complex<double> a, std::complex<double> b, bool complex_arith, double &c,
            std::complex<double> &s, std::complex<double> &r) {
  const double abs_a = complex_arith ? std::hypot(a.real(), a.imag()) : std::fabs(a.real());
  const double abs_b = complex_arith ? std::hypot(b.real(), b.imag()) : std::fabs(b.real());
  if (abs_b == 0.0) {
    c = 1.0;
    s = 0.0;
    r = a;
    return;
  }
  if (abs_a == 0.0) {
    c = 0.0;
    s = 1.0;
    r = b;
    return;
  }
  const double length = std::hypot(abs_a, abs_b);
  const std::complex<double> phase(a.real() / abs_a, a.imag() / abs_a);
  c = abs_a / length;
  const std::complex<double> cb(b.real() / length, -b.imag() / length);
  s = std::complex<double>(phase.real() * cb.real() - phase.imag() * cb.imag(),
                           phase.real() * cb.imag() + phase.imag() * cb.real());
  r = std::complex<double>(phase.real() * length, phase.imag() * length);
}

using cd = std::complex<double>;
// The reference's (num-complex) operations, written out so the host
// arithmetic does not depend on the C++ library's complex operators.
inline cd hmul(cd a, cd b) {
  return cd(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
}
inline cd hadd(cd a, cd b) { return cd(a.real() + b.real(), a.imag() + b.imag()); }
inline cd hsub(cd a, cd b) { return cd(a.real() - b.real(), a.imag() - b.imag()); }
inline cd hscale(cd a, double s) { return cd(a.real() * s, a.imag() * s); }
inline cd hconj(cd a) { return cd(a.real(), -a.imag()); }
inline cd hdiv(cd a, cd b) {
  const double n = b.real() * b.real() + b.imag() * b.imag();
  return cd((a.real() * b.real() + a.imag() * b.imag()) / n, (a.imag() * b.real() - a.real() * b.imag()) / n);
}
inline double hmod(cd a, bool complex_arith) {
  return complex_arith ? std::hypot(a.real(), a.imag()) : std::fabs(a.real());
}

}  // namespace

// ---------------------------------------------------------------------------
// DeviceGmres

fh_status DeviceGmres::reserve(size_t n, size_t restart) {
  if (n <= n_ && restart <= restart_ && basis_.ptr) return FH_OK;
  const size_t nn = std::max(n, n_), rr = std::max(restart, restart_);
  fh_status st;
  if ((st = basis_.allocate(nn * (rr + 1))) || (st = w_.allocate(nn)) || (st = tmp_.allocate(nn)) ||
      (st = residual_.allocate(nn)) || (st = col_.allocate(rr + 2)) || (st = y_.allocate(rr + 1)) ||
      (st = part_a_.allocate(kMaxBlocks)) || (st = part_b_.allocate(kMaxBlocks)) ||
      (st = mx_.allocate(kMaxBlocks)) || (st = ss_.allocate(kMaxBlocks)) || (st = scalars_.allocate(4)))
    return st;
  n_ = nn;
  restart_ = rr;
  return FH_OK;
}

size_t DeviceGmres::bytes() const {
  return n_ * (restart_ + 1) * 16 + 3 * n_ * 16 + (2 * restart_ + 3) * 16 + kMaxBlocks * 48 + 32;
}

// Scaled 2-norm of v into scalars_[slot].
fh_status DeviceGmres::norm(const double2 *d_v, size_t n, bool complex_arith, int slot,
                            hipStream_t stream) {
  const int nb = reduction_blocks(n);
  max_kernel<<<nb, kThreads, 0, stream>>>(d_v, n, complex_arith ? 1 : 0, mx_.ptr);
  sumsq_kernel<<<nb, kThreads, 0, stream>>>(d_v, n, complex_arith ? 1 : 0, mx_.ptr, nb, ss_.ptr);
  norm_finish_kernel<<<1, kThreads, 0, stream>>>(mx_.ptr, ss_.ptr, nb, scalars_.ptr + slot);
  FH_HIP_TRY(hipGetLastError());
  return FH_OK;
}

fh_status DeviceGmres::solve(const Op &apply, const double2 *d_diag, const double2 *d_b,
                             double2 *d_x, size_t n, const fh_gmres_params &params,
                             bool complex_arith, hipStream_t stream, GmresOutcome &out) {
  out = GmresOutcome{};
  if (n == 0) return FH_OK;
  fh_status st;
  size_t restart = params.restart < 1 ? 1 : params.restart;
  if (restart > n) restart = n;
  if ((st = reserve(n, restart))) return st;
  const int nb = reduction_blocks(n);
  const int cx = complex_arith ? 1 : 0;
  double host_scalar = 0.0;
  auto fetch = [&](int slot, double &v) -> fh_status {
    FH_HIP_TRY(hipMemcpyAsync(&host_scalar, scalars_.ptr + slot, sizeof(double), hipMemcpyDeviceToHost, stream));
    FH_HIP_TRY(hipStreamSynchronize(stream));
    v = host_scalar;
    return FH_OK;
  };
  // gmres.rs:100-110
  double b_norm;
  if ((st = norm(d_b, n, complex_arith, 0, stream)) || (st = fetch(0, b_norm))) return st;
  if (b_norm == 0.0) {
    zero2_kernel<<<grid_for(n), kThreads, 0, stream>>>(d_x, n);
    FH_HIP_TRY(hipGetLastError());
    FH_HIP_TRY(hipStreamSynchronize(stream));
    return FH_OK;
  }
  const double target = params.tolerance * b_norm;
  const size_t R1 = restart + 1;
  std::vector<cd> hess(R1 * restart), rs(restart), g(R1), y(restart);
  std::vector<double> rc(restart);
  std::vector<double2> colh(R1 + 1);
  size_t iterations = 0;
  auto residual_norm = [&](double &r_norm) -> fh_status {
    fh_status s2;
    if ((s2 = apply(d_x, residual_.ptr))) return s2;
    residual_kernel<<<grid_for(n), kThreads, 0, stream>>>(d_b, residual_.ptr, n);
    FH_HIP_TRY(hipGetLastError());
    if ((s2 = norm(residual_.ptr, n, complex_arith, 1, stream))) return s2;
    return fetch(1, r_norm);
  };
  double r_norm;
  if ((st = residual_norm(r_norm))) return st;
  for (;;) {
    if (r_norm <= target) {
      out.converged = true;
      break;
    }
    if (iterations >= params.max_iterations || !std::isfinite(r_norm)) {
      out.converged = false;
      break;
    }
    // basis_0 = r / |r|
    scale_kernel<<<grid_for(n), kThreads, 0, stream>>>(residual_.ptr, scalars_.ptr + 1, basis_.ptr, n);
    FH_HIP_TRY(hipGetLastError());
    size_t steps = 0, nbasis = 1;
    std::fill(g.begin(), g.end(), cd(0.0, 0.0));
    g[0] = cd(r_norm, 0.0);
    while (steps < restart && iterations < params.max_iterations) {
      const size_t j = steps;
      const double2 *vj = basis_.ptr + j * n;
      const double2 *pin = vj;
      if (d_diag) {
        precondition_kernel<<<grid_for(n), kThreads, 0, stream>>>(vj, d_diag, tmp_.ptr, n);
        FH_HIP_TRY(hipGetLastError());
        pin = tmp_.ptr;
      }
      if ((st = apply(pin, w_.ptr))) return st;
      ++iterations;
      zero2_kernel<<<1, kThreads, 0, stream>>>(col_.ptr, j + 2);
      // MGS twice over v_0 .. v_{nbasis-1}: launch t applies step t-1 and
      // starts step t (gmres.rs:147-154).
      const size_t total = 2 * nbasis;
      double2 *parts[2] = {part_a_.ptr, part_b_.ptr};
      for (size_t t = 0; t <= total; ++t) {
        const double2 *v_prev = t > 0 ? basis_.ptr + ((t - 1) % nbasis) * n : nullptr;
        const double2 *v_next = t < total ? basis_.ptr + (t % nbasis) * n : nullptr;
        mgs_kernel<<<nb, kThreads, 0, stream>>>(
            w_.ptr, v_prev, parts[(t + 1) & 1], nb, col_.ptr, t > 0 ? static_cast<int>((t - 1) % nbasis) : 0,
            v_next, parts[t & 1], mx_.ptr, n, cx);
      }
      FH_HIP_TRY(hipGetLastError());
      sumsq_kernel<<<nb, kThreads, 0, stream>>>(w_.ptr, n, cx, mx_.ptr, nb, ss_.ptr);
      norm_finish_kernel<<<1, kThreads, 0, stream>>>(mx_.ptr, ss_.ptr, nb, scalars_.ptr + 2);
      FH_HIP_TRY(hipGetLastError());
      FH_HIP_TRY(hipMemcpyAsync(colh.data(), col_.ptr, (j + 1) * sizeof(double2), hipMemcpyDeviceToHost, stream));
      FH_HIP_TRY(hipMemcpyAsync(&host_scalar, scalars_.ptr + 2, sizeof(double), hipMemcpyDeviceToHost, stream));
      FH_HIP_TRY(hipStreamSynchronize(stream));
      const double next_norm = host_scalar;
      cd *column = hess.data() + j * R1;
      for (size_t i = 0; i <= j; ++i) column[i] = cd(colh[i].x, colh[i].y);
      column[j + 1] = cd(next_norm, 0.0);
      for (size_t i = 0; i < j; ++i) {
        const cd a = column[i], bb = column[i + 1];
        column[i] = hadd(hscale(a, rc[i]), hmul(rs[i], bb));
        column[i + 1] = hsub(hscale(bb, rc[i]), hmul(hconj(rs[i]), a));
      }
      double c;
      cd s, r;
      givens(column[j], column[j + 1], complex_arith, c, s, r);
      column[j] = r;
      column[j + 1] = cd(0.0, 0.0);
      rc[j] = c;
      rs[j] = s;
      const cd gj = g[j];
      g[j] = hscale(gj, c);
      const cd t1 = hmul(hconj(s), gj);
      g[j + 1] = cd(-t1.real(), -t1.imag());
      ++steps;
      if (next_norm == 0.0 || hmod(g[j + 1], complex_arith) <= target) break;
      scale_kernel<<<grid_for(n), kThreads, 0, stream>>>(w_.ptr, scalars_.ptr + 2, basis_.ptr + nbasis * n, n);
      FH_HIP_TRY(hipGetLastError());
      ++nbasis;
    }
    // back substitution (gmres.rs:176-186)
    for (size_t ii = steps; ii-- > 0;) {
      cd sum = g[ii];
      for (size_t k = ii + 1; k < steps; ++k) sum = hsub(sum, hmul(hess[k * R1 + ii], y[k]));
      y[ii] = hdiv(sum, hess[ii * R1 + ii]);
    }
    std::vector<double2> yd(steps);
    for (size_t k = 0; k < steps; ++k) yd[k] = make_double2(y[k].real(), y[k].imag());
    if (steps) {
      FH_HIP_TRY(hipMemcpyAsync(y_.ptr, yd.data(), steps * sizeof(double2), hipMemcpyHostToDevice, stream));
      update_kernel<<<grid_for(n), kThreads, 0, stream>>>(basis_.ptr, y_.ptr, static_cast<int>(steps), d_diag,
                                                          d_x, n);
      FH_HIP_TRY(hipGetLastError());
    }
    if ((st = residual_norm(r_norm))) return st;
  }
  out.iterations = iterations;
  out.relative_residual = r_norm / b_norm;
  return FH_OK;
}

fh_status gmres_dense(size_t n, const double *a, const double *b, double *x,
                      const fh_gmres_params &params, bool complex_arith, GmresOutcome &out) {
  hipStream_t stream;
  FH_HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  struct Guard {
    hipStream_t s;
    ~Guard() { (void)hipStreamDestroy(s); }
  } guard{stream};
  DeviceArray<double2> da, db, dx;
  fh_status st;
  if ((st = da.allocate(n * n)) || (st = db.allocate(n)) || (st = dx.allocate(n))) return st;
  if (n) {
    FH_HIP_TRY(hipMemcpyAsync(da.ptr, a, n * n * sizeof(double2), hipMemcpyHostToDevice, stream));
    FH_HIP_TRY(hipMemcpyAsync(db.ptr, b, n * sizeof(double2), hipMemcpyHostToDevice, stream));
    FH_HIP_TRY(hipMemcpyAsync(dx.ptr, x, n * sizeof(double2), hipMemcpyHostToDevice, stream));
  }
  DeviceGmres solver;
  auto apply = [&](const double2 *in, double2 *o) -> fh_status {
    dense_apply_kernel<<<grid_for(n), kThreads, 0, stream>>>(da.ptr, in, o, n);
    FH_HIP_TRY(hipGetLastError());
    return FH_OK;
  };
  if ((st = solver.solve(apply, nullptr, db.ptr, dx.ptr, n, params, complex_arith, stream, out))) return st;
  if (n) FH_HIP_TRY(hipMemcpyAsync(x, dx.ptr, n * sizeof(double2), hipMemcpyDeviceToHost, stream));
  FH_HIP_TRY(hipStreamSynchronize(stream));
  return FH_OK;
}

// ---------------------------------------------------------------------------
// IterativeSystem

IterativeSystem::~IterativeSystem() {
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
}

size_t IterativeSystem::bytes() const {
  return op_.stats().memory_bytes + solver_.bytes() + n_ * 8 * 5 + m_ * 16 * 3;
}

fh_status IterativeSystem::build(const std::vector<fh_filament> &F, const std::vector<double> &res,
                                 const LoopBasis &basis, const fh_iterative_params &params,
                                 hipStream_t stream) {
  stream_ = stream;
  n_ = F.size();
  loops_ = basis.loops;
  m_ = basis.internal;
  p_ = loops_ - m_;
  gmres_ = params.gmres;
  fh_status st;
  FH_HIP_TRY(hipEventCreate(&ev0_));
  FH_HIP_TRY(hipEventCreate(&ev1_));
  const auto t0 = std::chrono::steady_clock::now();
  if ((st = op_.build(F, params.pfft, stream))) return st;
  operator_s_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  const size_t n = n_, m = m_, p = p_, loops = loops_;
  // device copies of the basis (rows), its transpose (branches), resistances
  const size_t nnz = basis.branch.size();
  std::vector<int64_t> tr_ptr(n + 1, 0);
  for (size_t i = 0; i < m; ++i)
    for (int64_t e = basis.row_ptr[i]; e < basis.row_ptr[i + 1]; ++e)
      tr_ptr[static_cast<size_t>(basis.branch[static_cast<size_t>(e)]) + 1]++;
  for (size_t b = 0; b < n; ++b) tr_ptr[b + 1] += tr_ptr[b];
  std::vector<int32_t> tr_loop(static_cast<size_t>(tr_ptr[n]));
  std::vector<double> tr_sign(static_cast<size_t>(tr_ptr[n]));
  {
    std::vector<int64_t> fill(tr_ptr.begin(), tr_ptr.end() - 1);
    for (size_t i = 0; i < m; ++i)
      for (int64_t e = basis.row_ptr[i]; e < basis.row_ptr[i + 1]; ++e) {
        const size_t b = static_cast<size_t>(basis.branch[static_cast<size_t>(e)]);
        const int64_t slot = fill[b]++;
        tr_loop[static_cast<size_t>(slot)] = static_cast<int32_t>(i);
        tr_sign[static_cast<size_t>(slot)] = basis.sign[static_cast<size_t>(e)];
      }
  }
  if ((st = d_res_.allocate(std::max<size_t>(n, 1))) || (st = d_row_ptr_.allocate(loops + 1)) ||
      (st = d_branch_.allocate(std::max<size_t>(nnz, 1))) || (st = d_sign_.allocate(std::max<size_t>(nnz, 1))) ||
      (st = d_tr_ptr_.allocate(n + 1)) || (st = d_tr_loop_.allocate(std::max<size_t>(tr_loop.size(), 1))) ||
      (st = d_tr_sign_.allocate(std::max<size_t>(tr_sign.size(), 1))) ||
      (st = d_cur_.allocate(std::max<size_t>(2 * n, 1))) || (st = d_flux_.allocate(std::max<size_t>(2 * n, 1))) ||
      (st = d_diag_.allocate(std::max<size_t>(m, 1))) || (st = d_rhs_.allocate(std::max<size_t>(m, 1))) ||
      (st = d_x_.allocate(std::max<size_t>(m, 1))))
    return st;
  if (n) FH_HIP_TRY(hipMemcpyAsync(d_res_.ptr, res.data(), n * sizeof(double), hipMemcpyHostToDevice, stream));
  FH_HIP_TRY(hipMemcpyAsync(d_row_ptr_.ptr, basis.row_ptr.data(), (loops + 1) * sizeof(int64_t), hipMemcpyHostToDevice, stream));
  if (nnz) {
    FH_HIP_TRY(hipMemcpyAsync(d_branch_.ptr, basis.branch.data(), nnz * sizeof(int32_t), hipMemcpyHostToDevice, stream));
    FH_HIP_TRY(hipMemcpyAsync(d_sign_.ptr, basis.sign.data(), nnz * sizeof(double), hipMemcpyHostToDevice, stream));
  }
  FH_HIP_TRY(hipMemcpyAsync(d_tr_ptr_.ptr, tr_ptr.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, stream));
  if (!tr_loop.empty()) {
    FH_HIP_TRY(hipMemcpyAsync(d_tr_loop_.ptr, tr_loop.data(), tr_loop.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream));
    FH_HIP_TRY(hipMemcpyAsync(d_tr_sign_.ptr, tr_sign.data(), tr_sign.size() * sizeof(double), hipMemcpyHostToDevice, stream));
  }
  // port columns (iterative.rs:185-209): one operator product per port,
  // all ports in one batched product
  r_ep_.assign(m * p, 0.0);
  l_ep_.assign(m * p, 0.0);
  r_pp_.assign(p * p, 0.0);
  l_pp_.assign(p * p, 0.0);
  if (p > 0 && n > 0) {
    std::vector<double> cur(p * n, 0.0), flux(p * n);
    for (size_t q = 0; q < p; ++q)
      for (int64_t e = basis.row_ptr[m + q]; e < basis.row_ptr[m + q + 1]; ++e)
        cur[q * n + static_cast<size_t>(basis.branch[static_cast<size_t>(e)])] += basis.sign[static_cast<size_t>(e)];
    DeviceArray<double> dc, df;
    if ((st = dc.allocate(p * n)) || (st = df.allocate(p * n))) return st;
    FH_HIP_TRY(hipMemcpyAsync(dc.ptr, cur.data(), p * n * sizeof(double), hipMemcpyHostToDevice, stream));
    if ((st = op_.apply(dc.ptr, df.ptr, static_cast<int>(p), stream))) return st;
    FH_HIP_TRY(hipMemcpyAsync(flux.data(), df.ptr, p * n * sizeof(double), hipMemcpyDeviceToHost, stream));
    FH_HIP_TRY(hipStreamSynchronize(stream));
    std::vector<double> drop(n);
    for (size_t q = 0; q < p; ++q) {
      for (size_t b = 0; b < n; ++b) drop[b] = cur[q * n + b] * res[b];
      for (size_t i = 0; i < loops; ++i) {
        double rs = 0.0, ls = 0.0;
        for (int64_t e = basis.row_ptr[i]; e < basis.row_ptr[i + 1]; ++e) {
          const size_t b = static_cast<size_t>(basis.branch[static_cast<size_t>(e)]);
          const double s = basis.sign[static_cast<size_t>(e)];
          rs += drop[b] * s;
          ls += flux[q * n + b] * s;
        }
        if (i < m) {
          r_ep_[i * p + q] = rs;
          l_ep_[i * p + q] = ls;
        } else {
          r_pp_[(i - m) * p + q] = rs;
          l_pp_[(i - m) * p + q] = ls;
        }
      }
    }
    for (std::vector<double> *X : {&r_pp_, &l_pp_}) {
      std::vector<double> t(p * p);
      for (size_t a = 0; a < p; ++a)
        for (size_t b = 0; b < p; ++b) t[a * p + b] = ((*X)[a * p + b] + (*X)[b * p + a]) * 0.5;
      *X = t;
    }
  }
  // Jacobi diagonals (iterative.rs:211-219, 520-536): the loop's self terms,
  // plus its mutual terms for loops of at most 32 branches, by the pair fill
  r_diag_.assign(m, 0.0);
  l_diag_.assign(m, 0.0);
  if (m > 0) {
    std::vector<uint2> pairs;
    std::vector<int64_t> first(m + 1, 0);
    for (size_t i = 0; i < m; ++i) {
      const int64_t b0 = basis.row_ptr[i], b1 = basis.row_ptr[i + 1];
      for (int64_t e = b0; e < b1; ++e) {
        const uint32_t a = static_cast<uint32_t>(basis.branch[static_cast<size_t>(e)]);
        pairs.push_back(make_uint2(a, a));
      }
      if (static_cast<size_t>(b1 - b0) <= kExactDiagonalMaxLoop)
        for (int64_t k = b0; k < b1; ++k)
          for (int64_t l = k + 1; l < b1; ++l)
            pairs.push_back(make_uint2(static_cast<uint32_t>(basis.branch[static_cast<size_t>(k)]),
                                       static_cast<uint32_t>(basis.branch[static_cast<size_t>(l)])));
      first[i + 1] = static_cast<int64_t>(pairs.size());
    }
    std::vector<double> vals(pairs.size());
    {
      DeviceArray<fh_filament> d_f;
      DeviceArray<uint2> d_pairs;
      DeviceArray<double> d_vals;
      if ((st = d_f.allocate(n)) || (st = d_pairs.allocate(pairs.size())) || (st = d_vals.allocate(pairs.size())))
        return st;
      FH_HIP_TRY(hipMemcpyAsync(d_f.ptr, F.data(), n * sizeof(fh_filament), hipMemcpyHostToDevice, stream));
      FH_HIP_TRY(hipMemcpyAsync(d_pairs.ptr, pairs.data(), pairs.size() * sizeof(uint2), hipMemcpyHostToDevice, stream));
      if ((st = fill_pairs(d_f.ptr, n, d_pairs.ptr, pairs.size(), d_vals.ptr, stream))) return st;
      FH_HIP_TRY(hipMemcpyAsync(vals.data(), d_vals.ptr, pairs.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
      FH_HIP_TRY(hipStreamSynchronize(stream));
    }
    for (size_t i = 0; i < m; ++i) {
      const int64_t b0 = basis.row_ptr[i], b1 = basis.row_ptr[i + 1];
      const size_t len = static_cast<size_t>(b1 - b0);
      double r = 0.0, own = 0.0;
      size_t q = static_cast<size_t>(first[i]);
      for (int64_t e = b0; e < b1; ++e) {
        r += res[static_cast<size_t>(basis.branch[static_cast<size_t>(e)])];
        own += vals[q++];
      }
      r_diag_[i] = r;
      if (len > kExactDiagonalMaxLoop) {
        l_diag_[i] = own;
        continue;
      }
      double mutual = 0.0;
      for (int64_t k = b0; k < b1; ++k)
        for (int64_t l = k + 1; l < b1; ++l)
          mutual += basis.sign[static_cast<size_t>(k)] * basis.sign[static_cast<size_t>(l)] * vals[q++];
      l_diag_[i] = own + 2.0 * mutual;
    }
  }
  return FH_OK;
}

fh_status IterativeSystem::apply_ee(double omega, bool dc, const double2 *x, double2 *y) {
  const size_t n = n_, m = m_;
  scatter_kernel<<<grid_for(n), kThreads, 0, stream_>>>(d_tr_ptr_.ptr, d_tr_loop_.ptr, d_tr_sign_.ptr, x, n,
                                                        d_cur_.ptr);
  FH_HIP_TRY(hipGetLastError());
  if (!dc) {
    if (profiling_) FH_HIP_TRY(hipEventRecord(ev0_, stream_));
    fh_status st = op_.apply(d_cur_.ptr, d_flux_.ptr, 2, stream_);
    if (st) return st;
    if (profiling_) {
      FH_HIP_TRY(hipEventRecord(ev1_, stream_));
      FH_HIP_TRY(hipEventSynchronize(ev1_));
      float ms = 0.0f;
      FH_HIP_TRY(hipEventElapsedTime(&ms, ev0_, ev1_));
      apply_ms_ += ms;
    }
    ++applies_;
  }
  gather_kernel<<<grid_for(m), kThreads, 0, stream_>>>(d_row_ptr_.ptr, d_branch_.ptr, d_sign_.ptr, d_res_.ptr,
                                                       d_cur_.ptr, d_flux_.ptr, omega, dc ? 1 : 0, n, m, y);
  FH_HIP_TRY(hipGetLastError());
  return FH_OK;
}

// iterative.rs:280-314 with schur_complement_gmres (:454-493).
fh_status IterativeSystem::impedance(double frequency, double *z, std::vector<size_t> &iterations,
                                     double &residual_max) {
  const size_t m = m_, p = p_;
  const bool dc = frequency == 0.0;
  const double omega = 6.283185307179586 * frequency;
  iterations.assign(p, 0);
  std::vector<cd> zz(p * p);
  for (size_t a = 0; a < p; ++a)
    for (size_t b = 0; b < p; ++b)
      zz[a * p + b] = cd(r_pp_[a * p + b], dc ? 0.0 : omega * l_pp_[a * p + b]);
  if (m > 0) {
    std::vector<double2> diag(m), rhs(m), solved(m * p);
    for (size_t i = 0; i < m; ++i) diag[i] = make_double2(r_diag_[i], dc ? 0.0 : omega * l_diag_[i]);
    FH_HIP_TRY(hipMemcpyAsync(d_diag_.ptr, diag.data(), m * sizeof(double2), hipMemcpyHostToDevice, stream_));
    auto apply = [&](const double2 *in, double2 *out) { return apply_ee(omega, dc, in, out); };
    for (size_t q = 0; q < p; ++q) {
      for (size_t i = 0; i < m; ++i) rhs[i] = make_double2(r_ep_[i * p + q], dc ? 0.0 : omega * l_ep_[i * p + q]);
      FH_HIP_TRY(hipMemcpyAsync(d_rhs_.ptr, rhs.data(), m * sizeof(double2), hipMemcpyHostToDevice, stream_));
      zero2_kernel<<<grid_for(m), kThreads, 0, stream_>>>(d_x_.ptr, m);
      FH_HIP_TRY(hipGetLastError());
      GmresOutcome o;
      fh_status st = solver_.solve(apply, d_diag_.ptr, d_rhs_.ptr, d_x_.ptr, m, gmres_, !dc, stream_, o);
      if (st) return st;
      iterations[q] = o.iterations;
      residual_max = std::max(residual_max, o.relative_residual);
      if (!o.converged)
        return set_error(FH_E_NOT_CONVERGED,
                         "GMRES did not converge at " + std::to_string(frequency) + " Hz for port " +
                             std::to_string(q) + ": relative residual " + std::to_string(o.relative_residual) +
                             " after " + std::to_string(o.iterations) + " iterations",
                         0, static_cast<int64_t>(q), o.relative_residual,
                         static_cast<int32_t>(std::min<size_t>(o.iterations, INT32_MAX)));
      FH_HIP_TRY(hipMemcpyAsync(rhs.data(), d_x_.ptr, m * sizeof(double2), hipMemcpyDeviceToHost, stream_));
      FH_HIP_TRY(hipStreamSynchronize(stream_));
      for (size_t i = 0; i < m; ++i) solved[i * p + q] = rhs[i];
    }
    // pp - ep^T solved
    for (size_t a = 0; a < p; ++a)
      for (size_t b = 0; b < p; ++b) {
        cd acc(0.0, 0.0);
        for (size_t i = 0; i < m; ++i) {
          const cd e(r_ep_[i * p + a], dc ? 0.0 : omega * l_ep_[i * p + a]);
          acc = hadd(acc, hmul(e, cd(solved[i * p + b].x, solved[i * p + b].y)));
        }
        zz[a * p + b] = hsub(zz[a * p + b], acc);
      }
  }
  for (size_t a = 0; a < p; ++a)
    for (size_t b = 0; b < p; ++b) {
      const cd e = hscale(hadd(zz[a * p + b], zz[b * p + a]), 0.5);
      z[2 * (a * p + b)] = e.real();
      z[2 * (a * p + b) + 1] = dc ? 0.0 : e.imag();
    }
  return FH_OK;
}

fh_status IterativeSystem::currents(double frequency, const double *excitation, std::vector<double> &internal,
                                    std::vector<double> &voltage, GmresOutcome &outcome) {
  const size_t m = m_, p = p_;
  const bool dc = frequency == 0.0;
  const double omega = 6.283185307179586 * frequency;
  auto ep = [&](size_t i, size_t a) { return cd(r_ep_[i * p + a], dc ? 0.0 : omega * l_ep_[i * p + a]); };
  std::vector<cd> ex(p), x(m);
  bool complex_ex = false;
  for (size_t a = 0; a < p; ++a) {
    ex[a] = cd(excitation[2 * a], excitation[2 * a + 1]);
    complex_ex = complex_ex || ex[a].imag() != 0.0;
  }
  outcome = GmresOutcome{};
  if (m > 0) {
    std::vector<double2> diag(m), rhs(m);
    for (size_t i = 0; i < m; ++i) {
      diag[i] = make_double2(r_diag_[i], dc ? 0.0 : omega * l_diag_[i]);
      cd acc(0.0, 0.0);
      for (size_t a = 0; a < p; ++a) acc = hadd(acc, hmul(ep(i, a), ex[a]));
      rhs[i] = make_double2(acc.real(), acc.imag());
    }
    FH_HIP_TRY(hipMemcpyAsync(d_diag_.ptr, diag.data(), m * sizeof(double2), hipMemcpyHostToDevice, stream_));
    FH_HIP_TRY(hipMemcpyAsync(d_rhs_.ptr, rhs.data(), m * sizeof(double2), hipMemcpyHostToDevice, stream_));
    zero2_kernel<<<grid_for(m), kThreads, 0, stream_>>>(d_x_.ptr, m);
    FH_HIP_TRY(hipGetLastError());
    auto apply = [&](const double2 *in, double2 *out) { return apply_ee(omega, dc, in, out); };
    // the real form at DC for a real excitation (every imaginary part zero)
    fh_status st = solver_.solve(apply, d_diag_.ptr, d_rhs_.ptr, d_x_.ptr, m, gmres_, !dc || complex_ex, stream_,
                                 outcome);
    if (st) return st;
    if (!outcome.converged)
      return set_error(FH_E_NOT_CONVERGED,
                       "GMRES did not converge at " + std::to_string(frequency) + " Hz: relative residual " +
                           std::to_string(outcome.relative_residual) + " after " +
                           std::to_string(outcome.iterations) + " iterations",
                       0, 0, outcome.relative_residual,
                       static_cast<int32_t>(std::min<size_t>(outcome.iterations, INT32_MAX)));
    FH_HIP_TRY(hipMemcpyAsync(rhs.data(), d_x_.ptr, m * sizeof(double2), hipMemcpyDeviceToHost, stream_));
    FH_HIP_TRY(hipStreamSynchronize(stream_));
    for (size_t i = 0; i < m; ++i) x[i] = cd(-rhs[i].x, -rhs[i].y);
  }
  internal.assign(2 * m, 0.0);
  for (size_t i = 0; i < m; ++i) {
    internal[2 * i] = x[i].real();
    internal[2 * i + 1] = x[i].imag();
  }
  voltage.assign(2 * p, 0.0);
  for (size_t a = 0; a < p; ++a) {
    cd acc(0.0, 0.0);
    for (size_t b = 0; b < p; ++b)
      acc = hadd(acc, hmul(cd(r_pp_[a * p + b], dc ? 0.0 : omega * l_pp_[a * p + b]), ex[b]));
    cd tr(0.0, 0.0);
    for (size_t i = 0; i < m; ++i) tr = hadd(tr, hmul(ep(i, a), x[i]));
    acc = hadd(acc, tr);
    voltage[2 * a] = acc.real();
    voltage[2 * a + 1] = acc.imag();
  }
  return FH_OK;
}

}  // namespace fh
