// Per-frequency solve of the mesh system on gfx950: replaces lu_solve
// (fasterhenry/src/dense.rs:43-189) and schur_complement
// (fasterhenry/src/solve.rs:1154-1164) for MeshSystem::impedance
// (solve.rs:806-830).
//
// Z(w) = [[Zee, Zep], [Zep^T, Zpp]] = M (R + jwL) M^T is complex symmetric
// with real and imaginary parts positive definite (R and L are), so the
// symmetric factorization Zee = L D L^T needs no pivoting: its growth factor
// is bounded by 3 (N. J. Higham, "Factorizing complex symmetric matrices
// with positive definite real and imaginary parts", Math. Comp. 67 (1998)).
// Eliminating the m internal loops of the bordered (m + p) x (m + p) matrix
// leaves Zpp - Zep^T Zee^-1 Zep = Z in its trailing p x p block: the Schur
// complement the reference forms with an LU solve and a product.
//
// All frequencies of a sweep chunk are factorized together: each step is
// one launch over (frequency, tile), so the latency-bound diagonal block and
// panel of one frequency overlap with the others.
//
// Storage: planar complex (re plane, im plane), column-major, leading
// dimension ld (a multiple of 64), one ld x ld matrix per frequency; only the
// lower triangle is referenced.
//
// Steps for the block column [k0, k0 + w):
//  1. diag_kernel   -- unblocked L D L^T of the w x w diagonal block in LDS;
//  2. panel_kernel  -- W21 = A21 L11^-T (row-wise forward substitution),
//                      L21 = W21 D^-1, both stored;
//  3. update_kernel -- A22 -= W21 L21^T on the lower triangle (the O(m^3)
//                      part), 64 x 64 output tiles, 4 x 4 complex
//                      micro-tiles per thread, explicit FMA.
#include <hip/hip_runtime.h>

#include <cmath>

#include "fh_ldlt.hpp"

namespace fh {
namespace {

constexpr int kNb = 32;    // block column width
constexpr int kUT = 64;    // update output tile
constexpr int kUK = 16;    // update k-chunk staged in LDS

template <typename T>
struct Cx {
  T re, im;
};

template <typename T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> a, Cx<T> b) {
  return {__builtin_fma(a.re, b.re, -(a.im * b.im)), __builtin_fma(a.re, b.im, a.im * b.re)};
}

// 1 / d by the textbook formula (conj(d) / |d|^2).
template <typename T>
__device__ __forceinline__ Cx<T> crecip(Cx<T> d) {
  const T n = d.re * d.re + d.im * d.im;
  return {d.re / n, -d.im / n};
}

// ---------------------------------------------------------------------------
// Forming Z(w): A = R + j w L on the lower triangle.

template <typename T>
__global__ void form_kernel(const double *__restrict__ lmesh, int64_t loops,
                            const double *__restrict__ omegas, int N, int64_t ld,
                            T *__restrict__ re, T *__restrict__ im, int64_t mat_stride) {
  const int f = blockIdx.y;
  const double omega = omegas[f];
  T *Re = re + f * mat_stride, *Im = im + f * mat_stride;
  const int64_t total = static_cast<int64_t>(N) * ld;
  for (int64_t idx = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; idx < total;
       idx += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t c = idx / ld, r = idx % ld;
    if (r >= N || r < c) {
      Re[idx] = T(0);
      Im[idx] = T(0);
      continue;
    }
    Re[idx] = T(0);
    Im[idx] = static_cast<T>(omega * lmesh[r * loops + c]);
  }
}

template <typename T>
__global__ void scatter_r_kernel(const int32_t *__restrict__ rows, const int32_t *__restrict__ cols,
                                 const double *__restrict__ vals, int64_t nnz, int64_t ld,
                                 T *__restrict__ re, int64_t mat_stride) {
  const int f = blockIdx.y;
  T *Re = re + f * mat_stride;
  for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < nnz;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int32_t r = rows[e], c = cols[e];
    if (r >= c) Re[static_cast<int64_t>(c) * ld + r] = static_cast<T>(vals[e]);
  }
}

// ---------------------------------------------------------------------------
// 1. Diagonal block.

template <typename T>
__global__ void __launch_bounds__(256) diag_kernel(T *__restrict__ re, T *__restrict__ im,
                                                   int64_t ld, int64_t mat_stride, int k0, int w,
                                                   int *__restrict__ singular) {
  const int f = blockIdx.x;
  T *Re = re + f * mat_stride, *Im = im + f * mat_stride;
  __shared__ T sr[kNb][kNb + 1], si[kNb][kNb + 1];
  for (int q = threadIdx.x; q < w * w; q += blockDim.x) {
    const int c = q / w, r = q % w;
    sr[r][c] = Re[(k0 + c) * ld + k0 + r];
    si[r][c] = Im[(k0 + c) * ld + k0 + r];
  }
  __syncthreads();
  for (int c = 0; c < w; ++c) {
    const Cx<T> d{sr[c][c], si[c][c]};
    if (d.re == T(0) && d.im == T(0)) {
      if (threadIdx.x == 0) singular[f] = 1;
    }
    const Cx<T> inv = crecip(d);
    __syncthreads();
    // trailing part of the block: A[r][j] -= l_rc * w_jc, c < j <= r
    const int rem = w - c - 1;
    for (int q = threadIdx.x; q < rem * rem; q += blockDim.x) {
      const int r = c + 1 + q / rem, j = c + 1 + q % rem;
      if (j > r) continue;
      const Cx<T> wr{sr[r][c], si[r][c]}, wj{sr[j][c], si[j][c]};
      const Cx<T> l = cmul(wr, inv);
      const Cx<T> prod = cmul(l, wj);
      sr[r][j] -= prod.re;
      si[r][j] -= prod.im;
    }
    __syncthreads();
    for (int r = c + 1 + threadIdx.x; r < w; r += blockDim.x) {
      const Cx<T> wr{sr[r][c], si[r][c]};
      const Cx<T> l = cmul(wr, inv);
      sr[r][c] = l.re;
      si[r][c] = l.im;
    }
    __syncthreads();
  }
  for (int q = threadIdx.x; q < w * w; q += blockDim.x) {
    const int c = q / w, r = q % w;
    if (r >= c) {
      Re[(k0 + c) * ld + k0 + r] = sr[r][c];
      Im[(k0 + c) * ld + k0 + r] = si[r][c];
    }
  }
}

// ---------------------------------------------------------------------------
// 2. Panel: rows r >= k0 + w.  W[r][c] = A[r][c] - sum_{j<c} W[r][j] L11[c][j],
// L21[r][c] = W[r][c] / D[c].

template <typename T>
__global__ void __launch_bounds__(256) panel_kernel(T *__restrict__ re, T *__restrict__ im,
                                                    int64_t ld, int64_t mat_stride, int k0, int w,
                                                    int N, T *__restrict__ wre, T *__restrict__ wim,
                                                    int64_t w_stride) {
  const int f = blockIdx.y;
  T *Re = re + f * mat_stride, *Im = im + f * mat_stride;
  T *Wr = wre + f * w_stride, *Wi = wim + f * w_stride;
  __shared__ T lr[kNb][kNb + 1], li[kNb][kNb + 1];
  for (int q = threadIdx.x; q < w * w; q += blockDim.x) {
    const int c = q / w, r = q % w;
    lr[r][c] = Re[(k0 + c) * ld + k0 + r];
    li[r][c] = Im[(k0 + c) * ld + k0 + r];
  }
  __syncthreads();
  const int64_t row = static_cast<int64_t>(k0) + w + blockIdx.x * static_cast<int64_t>(blockDim.x) +
                      threadIdx.x;
  if (row >= N) return;
  Cx<T> wv[kNb];
  for (int c = 0; c < w; ++c) {
    Cx<T> a{Re[(k0 + c) * ld + row], Im[(k0 + c) * ld + row]};
    for (int j = 0; j < c; ++j) {
      const Cx<T> p = cmul(wv[j], Cx<T>{lr[c][j], li[c][j]});
      a.re -= p.re;
      a.im -= p.im;
    }
    wv[c] = a;
  }
  for (int c = 0; c < w; ++c) {
    const Cx<T> inv = crecip(Cx<T>{lr[c][c], li[c][c]});
    const Cx<T> l = cmul(wv[c], inv);
    Re[(k0 + c) * ld + row] = l.re;
    Im[(k0 + c) * ld + row] = l.im;
    Wr[c * ld + row] = wv[c].re;
    Wi[c * ld + row] = wv[c].im;
  }
}

// ---------------------------------------------------------------------------
// 3. Trailing update of the lower triangle: A[r][s] -= sum_c W[r][c] L[s][c],
// rows and columns >= s0 = k0 + w.  Tile (tr, tc) with tr >= tc.

template <typename T>
__global__ void __launch_bounds__(256) update_kernel(T *__restrict__ re, T *__restrict__ im,
                                                     int64_t ld, int64_t mat_stride, int k0,
                                                     int w, int N, const T *__restrict__ wre,
                                                     const T *__restrict__ wim, int64_t w_stride,
                                                     int ntiles) {
  const int f = blockIdx.y;
  // lower-triangular tile index -> (tr, tc), tr >= tc
  const int t = blockIdx.x;
  int tr = static_cast<int>((__builtin_sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while (tr * (tr + 1) / 2 > t) --tr;
  while ((tr + 1) * (tr + 2) / 2 <= t) ++tr;
  const int tc = t - tr * (tr + 1) / 2;
  if (tr >= ntiles) return;
  T *Re = re + f * mat_stride, *Im = im + f * mat_stride;
  const T *Wr = wre + f * w_stride, *Wi = wim + f * w_stride;
  const int s0 = k0 + w;
  const int r0 = s0 + tr * kUT, c0 = s0 + tc * kUT;
  __shared__ T aw_r[kUK][kUT], aw_i[kUK][kUT], bl_r[kUK][kUT], bl_i[kUK][kUT];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  Cx<T> acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = {T(0), T(0)};
  for (int kc = 0; kc < w; kc += kUK) {
    __syncthreads();
    for (int q = threadIdx.x; q < kUK * kUT; q += blockDim.x) {
      const int k = q / kUT, x = q % kUT;
      const int kk = kc + k;
      const int64_t rr = r0 + x, cc = c0 + x;
      const bool kin = kk < w;
      aw_r[k][x] = (kin && rr < N) ? Wr[kk * ld + rr] : T(0);
      aw_i[k][x] = (kin && rr < N) ? Wi[kk * ld + rr] : T(0);
      bl_r[k][x] = (kin && cc < N) ? Re[(k0 + kk) * ld + cc] : T(0);
      bl_i[k][x] = (kin && cc < N) ? Im[(k0 + kk) * ld + cc] : T(0);
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < kUK; ++k) {
      Cx<T> av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) av[a] = {aw_r[k][ty + 16 * a], aw_i[k][ty + 16 * a]};
#pragma unroll
      for (int b = 0; b < 4; ++b) bv[b] = {bl_r[k][tx + 16 * b], bl_i[k][tx + 16 * b]};
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          acc[a][b].re = __builtin_fma(av[a].re, bv[b].re, acc[a][b].re);
          acc[a][b].re = __builtin_fma(-av[a].im, bv[b].im, acc[a][b].re);
          acc[a][b].im = __builtin_fma(av[a].re, bv[b].im, acc[a][b].im);
          acc[a][b].im = __builtin_fma(av[a].im, bv[b].re, acc[a][b].im);
        }
    }
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int64_t r = r0 + ty + 16 * a;
    if (r >= N) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int64_t c = c0 + tx + 16 * b;
      if (c >= N || c > r) continue;
      Re[c * ld + r] -= acc[a][b].re;
      Im[c * ld + r] -= acc[a][b].im;
    }
  }
}

// Z_f (p x p) from the trailing block; f = 0 gets Im Z = 0 exactly
// (solve.rs:809-812 solves DC in real arithmetic).
template <typename T>
__global__ void extract_kernel(const T *__restrict__ re, const T *__restrict__ im, int64_t ld,
                               int64_t mat_stride, int m, int p, const double *__restrict__ omegas,
                               double *__restrict__ z) {
  const int f = blockIdx.x;
  const T *Re = re + f * mat_stride, *Im = im + f * mat_stride;
  for (int q = threadIdx.x; q < p * p; q += blockDim.x) {
    const int i = q / p, j = q % p;
    const int r = m + (i > j ? i : j), c = m + (i > j ? j : i);
    z[(static_cast<int64_t>(f) * p * p + q) * 2] = static_cast<double>(Re[c * ld + r]);
    z[(static_cast<int64_t>(f) * p * p + q) * 2 + 1] =
        omegas[f] == 0.0 ? 0.0 : static_cast<double>(Im[c * ld + r]);
  }
}

}  // namespace

template <typename T>
fh_status ldlt_batch_t(const DenseSolveInputs &in, const double *d_omegas, int K, double *d_z,
                       int *d_singular, LdltWorkspace &ws, hipStream_t stream, Profiler *prof) {
  const int m = static_cast<int>(in.m), p = static_cast<int>(in.p), N = m + p;
  const int64_t ld = ((static_cast<int64_t>(N) + 63) / 64) * 64;
  const int64_t mat_stride = ld * ld;
  const int64_t w_stride = ld * kNb;
  fh_status st;
  const size_t need = static_cast<size_t>(K) * mat_stride * 2 * sizeof(T) +
                      static_cast<size_t>(K) * w_stride * 2 * sizeof(T);
  if ((st = ws.reserve(need))) return st;
  T *re = reinterpret_cast<T *>(ws.base);
  T *im = re + static_cast<size_t>(K) * mat_stride;
  T *wre = im + static_cast<size_t>(K) * mat_stride;
  T *wim = wre + static_cast<size_t>(K) * w_stride;
  {
    dim3 grid(1024, K);
    form_kernel<T><<<grid, 256, 0, stream>>>(in.d_lmesh, static_cast<int64_t>(N), d_omegas, N, ld,
                                             re, im, mat_stride);
    FH_HIP_TRY(hipGetLastError());
    if (in.r_nnz) {
      dim3 g2(static_cast<unsigned>((in.r_nnz + 255) / 256 < 4096 ? (in.r_nnz + 255) / 256 : 4096), K);
      scatter_r_kernel<T><<<g2, 256, 0, stream>>>(in.d_r_rows, in.d_r_cols, in.d_r_vals,
                                                  static_cast<int64_t>(in.r_nnz), ld, re, mat_stride);
      FH_HIP_TRY(hipGetLastError());
    }
  }
  for (int k0 = 0; k0 < m; k0 += kNb) {
    const int w = m - k0 < kNb ? m - k0 : kNb;
    diag_kernel<T><<<K, 256, 0, stream>>>(re, im, ld, mat_stride, k0, w, d_singular);
    FH_HIP_TRY(hipGetLastError());
    const int rows = N - (k0 + w);
    if (rows <= 0) continue;
    dim3 pg(static_cast<unsigned>((rows + 255) / 256), K);
    panel_kernel<T><<<pg, 256, 0, stream>>>(re, im, ld, mat_stride, k0, w, N, wre, wim, w_stride);
    FH_HIP_TRY(hipGetLastError());
    const int ntiles = (rows + kUT - 1) / kUT;
    dim3 ug(static_cast<unsigned>(ntiles * (ntiles + 1) / 2), K);
    if (prof) prof->begin(stream);
    update_kernel<T><<<ug, 256, 0, stream>>>(re, im, ld, mat_stride, k0, w, N, wre, wim, w_stride,
                                             ntiles);
    FH_HIP_TRY(hipGetLastError());
    if (prof) {
      const double t = static_cast<double>(rows);
      prof->end(stream, 8.0 * w * (t * (t + 1.0) / 2.0) * K);
    }
  }
  extract_kernel<T><<<K, 64, 0, stream>>>(re, im, ld, mat_stride, m, p, d_omegas, d_z);
  FH_HIP_TRY(hipGetLastError());
  return FH_OK;
}

fh_status ldlt_solve_batch(const DenseSolveInputs &in, const double *d_omegas, int K, double *d_z,
                           int *d_singular, LdltWorkspace &ws, hipStream_t stream, Profiler *prof) {
  return ldlt_batch_t<double>(in, d_omegas, K, d_z, d_singular, ws, stream, prof);
}

size_t ldlt_bytes_per_frequency(size_t m, size_t p) {
  const size_t N = m + p;
  const size_t ld = ((N + 63) / 64) * 64;
  return ld * ld * 2 * sizeof(double) + ld * kNb * 2 * sizeof(double);
}

}  // namespace fh
