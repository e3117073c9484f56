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
//
// Plain FP64 form: eliminating the m internal loops of the bordered
// (m + p) x (m + p) matrix leaves Zpp - Zep^T Zee^-1 Zep = Z in its trailing
// p x p block -- the Schur complement the reference forms with an LU solve
// and a product.
//
// Refined form: the same factorization in FP32, then FP64 iterative
// refinement of x = Zee^-1 Zep (residuals with the FP64 R and L, solves with
// the FP32 factors), and Z = Zpp - Zep^T x - x^T r, where r = Zep - Zee x is
// the last residual: the correction makes the error in Z second order in the
// error of x.
//
// All frequencies of a sweep chunk are factorized together: each step is one
// launch over (tile, frequency), so the latency-bound diagonal block of one
// frequency overlaps with the others and launch counts do not grow with K.
//
// Storage: planar complex (re plane, im plane), column-major, leading
// dimension ld (a multiple of 64), one ld x ld matrix per frequency; only the
// lower triangle is referenced.  Two-level right-looking blocking: outer
// blocks of NB columns (a multiple of 64), each factorized 64 columns at a
// time:
//  1. diag_kernel   -- L D L^T of the 64 x 64 diagonal block in LDS, and the
//                      inverse of its unit-lower factor (kept for the solves);
//  2. panel_kernel  -- W21 = A21 L11^-T as a GEMM with that inverse (MFMA),
//                      L21 = W21 D^-1, for all rows below;
//  3. update_kernel -- inside the outer block: the columns still to be
//                      factorized; after it: A22 -= W L^T over the NB columns
//                      (the O(m^3) part), so the read-modify-write of A22
//                      costs one pass per NB columns instead of per 64.
// The update is a 64 x 64 output tile per workgroup, four waves of 2 x 2
// 16x16x4 MFMA blocks, complex as four real products.
//
// The FP32 factorization of the refined form replaces step 3's outer update
// by a left-looking one on the bf16 matrix cores (update_left_kernel): each
// outer block first receives the whole contribution of the columns to its
// left in one pass, from split-bf16 copies of L and W written once per block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <cmath>
#include <type_traits>

#include "fh_ldlt.hpp"

namespace fh {
namespace {

constexpr int kB = 64;        // inner block (diagonal block, panel width)
constexpr int kT = 64;        // GEMM output tile (both dimensions)
constexpr int kKC = 16;       // k-chunk staged in LDS
constexpr int kNbPlain = 128; // outer block, FP64 form
constexpr int kNbRefined = 256;  // outer block, FP32 form
constexpr int kMaxRefine = 30;
constexpr int kMaxGmresRefine = 8;   // outer steps under GMRES-IR
constexpr int kGmresRestart = 20;    // inner GMRES steps per outer step
constexpr double kGmresTol = 1e-8;   // inner reduction of the preconditioned residual
constexpr int kGraphSteps = 2;       // classical steps inside the first factorization's graph

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// The 16x16x4 instructions.  A and B operands: lane l holds A[l % 16][l / 16]
// and B[l / 16][l % 16].  Accumulator register v of lane l holds
// D[row(l, v)][l % 16]: rows 4 (l / 16) + v for FP32, l / 16 + 4 v for FP64.
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  using acc = d4;
  __device__ static acc run(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <>
struct Mfma<float> {
  using acc = f4;
  __device__ static acc run(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

struct Z2 {
  double re, im;
};
__device__ __forceinline__ Z2 zmul(Z2 a, Z2 b) {
  return {__builtin_fma(a.re, b.re, -(a.im * b.im)), __builtin_fma(a.re, b.im, a.im * b.re)};
}
__device__ __forceinline__ Z2 zrecip(Z2 d) {
  const double n = d.re * d.re + d.im * d.im;
  return {d.re / n, -d.im / n};
}

// bf16 splitting for the matrix-core products of the FP32 factorization
// (see the left-looking update below): x = hi + lo in bf16.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ void split_bf16(float x, __bf16 &hi, __bf16 &lo) {
  hi = static_cast<__bf16>(x);
  lo = static_cast<__bf16>(x - static_cast<float>(hi));
}

typedef int i32x4 __attribute__((ext_vector_type(4)));

// -v: the sign bits of the eight bf16, one 32-bit XOR per pair
__device__ __forceinline__ bf16x8 neg8(bf16x8 v) {
  i32x4 s = __builtin_bit_cast(i32x4, v);
  s ^= static_cast<int>(0x80008000u);
  return __builtin_bit_cast(bf16x8, s);
}

// Block index of the split operand copies (see the left-looking update).
__host__ __device__ __forceinline__ int64_t split_block(int kc, int g, int G) {
  return static_cast<int64_t>(kc) * G - static_cast<int64_t>(kc) * (kc - 1) + (g - 2 * kc);
}

// Left-looking update of one strip, done by one wave inside the kernel that
// next reads it: A[r][s] -= sum over the chunks kc in [kc_beg, kc_end) of
// W[r][k] L[s][k], for the 64 rows [rb, rb + 64) (r < rend) and the 16
// columns [cb, cb + 16) (s < cend, s <= r), from the split copies.  The
// products, their order and the final subtraction are those of
// update_left_kernel for the same entries, so the bits are the same.
__device__ __forceinline__ void left_strip_update(float *__restrict__ Re, float *__restrict__ Im,
                                                  int64_t ld, int rb, int rend, int cb, int cend,
                                                  int kc_beg, int kc_end, const bf16x8 *lsplit,
                                                  const bf16x8 *wsplit, int64_t split_stride,
                                                  int G, int lane) {
  const __amdgpu_buffer_rsrc_t rL = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<bf16x8 *>(lsplit), 0, static_cast<int>(split_stride * 16), 0x00020000);
  const __amdgpu_buffer_rsrc_t rW = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<bf16x8 *>(wsplit), 0, static_cast<int>(split_stride * 16), 0x00020000);
  const int voff = lane * 16;
  const int gs = min(cb / 16, G - 1);
  int gr[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) gr[b] = min(rb / 16 + b, G - 1);
  f4 accr[4], acci[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    accr[b] = f4{0, 0, 0, 0};
    acci[b] = f4{0, 0, 0, 0};
  }
  for (int kc = kc_beg; kc < kc_end; ++kc) {
    bf16x8 p[4], q[4][4];
    const int so = static_cast<int>(split_block(kc, gs, G) * 4096);
#pragma unroll
    for (int pl = 0; pl < 4; ++pl)
      p[pl] = __builtin_bit_cast(bf16x8,
                                 __builtin_amdgcn_raw_buffer_load_b128(rL, voff + 1024 * pl, so, 0));
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int sq = static_cast<int>(split_block(kc, gr[b], G) * 4096);
#pragma unroll
      for (int pl = 0; pl < 4; ++pl)
        q[b][pl] = __builtin_bit_cast(
            bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rW, voff + 1024 * pl, sq, 0));
    }
    const bf16x8 nih = neg8(p[2]), nil = neg8(p[3]);
#pragma unroll
    for (int jb = 0; jb < 4; ++jb) {
      f4 cr = accr[jb], ci = acci[jb];
      cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[0], q[jb][0], cr, 0, 0, 0);
      cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[0], q[jb][1], cr, 0, 0, 0);
      cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[1], q[jb][0], cr, 0, 0, 0);
      cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nih, q[jb][2], cr, 0, 0, 0);
      cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nih, q[jb][3], cr, 0, 0, 0);
      cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nil, q[jb][2], cr, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[0], q[jb][2], ci, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[0], q[jb][3], ci, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[1], q[jb][2], ci, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[2], q[jb][0], ci, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[2], q[jb][1], ci, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[3], q[jb][0], ci, 0, 0, 0);
      accr[jb] = cr;
      acci[jb] = ci;
    }
  }
  const int kq = lane >> 4, x16 = lane & 15;
  float cr[4][4], ci[4][4];
#pragma unroll
  for (int jb = 0; jb < 4; ++jb)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t s = min(cb + 4 * kq + v, cend - 1);
      const int64_t r = min(rb + 16 * jb + x16, rend - 1);
      cr[jb][v] = Re[s * ld + r];
      ci[jb][v] = Im[s * ld + r];
    }
#pragma unroll
  for (int jb = 0; jb < 4; ++jb)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t s = cb + 4 * kq + v;
      const int64_t r = rb + 16 * jb + x16;
      if (s <= r && r < rend && s < cend) {
        Re[s * ld + r] = cr[jb][v] - accr[jb][v];
        Im[s * ld + r] = ci[jb][v] - acci[jb][v];
      }
    }
}

// ---------------------------------------------------------------------------
// Forming Z(w) = R + j w L on the lower triangle (solve.rs:1003-1009), all
// frequencies of the batch per thread: l_mesh is read once.  l_mesh is
// symmetric, so its row c is read for column c of A (row-contiguous).  Only
// the lower triangle (r >= c, r < N) is written; no kernel reads the rest.
// Each thread writes one 16-byte chunk of a column (rows [V q, V q + V),
// V = 16 / sizeof(T)) of every frequency's matrix: ld is a multiple of 64,
// so chunks are aligned, and a chunk that meets the diagonal or the end of
// the column is written element by element.  Row y of the grid covers the
// chunks of column y followed by those of column N - 1 - y: about N / V + 2
// chunks for every y, no idle workgroups.

template <typename T>
__global__ void form_kernel(const double *__restrict__ lmesh, int64_t loops,
                            const double *__restrict__ omegas, int K, int N, int64_t ld,
                            T *__restrict__ re, T *__restrict__ im, int64_t mat_stride, int cend) {
  constexpr int V = 16 / static_cast<int>(sizeof(T));
  typedef T TV __attribute__((ext_vector_type(V)));
  const int y = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nch = (N + V - 1) / V;
  const int n1 = nch - y / V, c2 = N - 1 - y;
  const bool first = i < n1;
  // the last row of an odd N pairs its column with itself: once
  if (!first && c2 <= y) return;
  const int c = first ? y : c2;
  if (c >= cend) return;  // columns the trailing update forms itself
  const int chunk = first ? y / V + i : c2 / V + (i - n1);
  if (chunk >= nch) return;
  const int r0 = chunk * V;
  double l[V];
  bool ok[V];
#pragma unroll
  for (int e = 0; e < V; ++e) {
    const int r = r0 + e;
    ok[e] = r >= c && r < N;
    l[e] = ok[e] ? lmesh[static_cast<int64_t>(c) * loops + r] : 0.0;
  }
  const int64_t a = static_cast<int64_t>(c) * ld + r0;
  const bool full = r0 >= c && r0 + V <= N;
  for (int f = 0; f < K; ++f) {
    const double w = omegas[f];
    T *pr = re + f * mat_stride + a, *pi = im + f * mat_stride + a;
    if (full) {
      TV vi;
#pragma unroll
      for (int e = 0; e < V; ++e) vi[e] = static_cast<T>(w * l[e]);
      *reinterpret_cast<TV *>(pr) = TV(0);
      *reinterpret_cast<TV *>(pi) = vi;
    } else {
#pragma unroll
      for (int e = 0; e < V; ++e)
        if (ok[e]) {
          pr[e] = T(0);
          pi[e] = static_cast<T>(w * l[e]);
        }
    }
  }
}

template <typename T>
__global__ void scatter_r_kernel(const int32_t *__restrict__ rows, const int32_t *__restrict__ cols,
                                 const double *__restrict__ vals, int64_t nnz, int64_t ld,
                                 T *__restrict__ re, int64_t mat_stride, int cend) {
  T *Re = re + blockIdx.y * mat_stride;
  for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < nnz;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int32_t r = rows[e], c = cols[e];
    if (r >= c && c < cend) Re[static_cast<int64_t>(c) * ld + r] = static_cast<T>(vals[e]);
  }
}

// ---------------------------------------------------------------------------
// 1. Diagonal block at k0 (width w <= 64, block index blk): A11 = L11 D L11^T,
// then L11^-1.  Writes L11 and D in place, linvt[k][c] = Linv[c][k] (zero
// outside the block) and D^-1 for block blk of frequency f.
//
// One workgroup per frequency; lane r of every wave holds row r, and wave g
// of the W waves holds the columns S g .. S g + S - 1 in registers (S = 64 / W
// slots), so a column step costs each lane S complex multiply-adds and one
// barrier.  Pivot columns
// reach the other waves through a two-entry LDS ring.  Linv is built in the
// same pass: the row operations of the elimination (row r -= l_rc row c) are
// applied to an identity X, which becomes L11^-1.  Slot j holds A[r][j]
// until column j is eliminated and X[r][j] from then on (X[r][j] is the
// identity until that step), so each slot has one update: for j > c,
// A[r][j] -= l_r A[j][c] with A[j][c] from the ring; for j < c,
// X[r][j] -= l_r X[c][j] with X[c][j] read from lane c; for j = c the slot
// becomes X[r][c] = delta_rc - l_r.
//
// Column c = S q + t is held by wave q, slot t.  The step loop is unrolled
// over t, so in every wave each slot's role at a step is fixed by one
// wave-uniform branch: waves below q hold only columns left of c (all X
// updates), waves above q only columns right of it (all A updates), and in
// wave q the roles split at slot t, known at compile time.  (A round-robin
// column assignment makes every slot's role a run-time choice, and the
// selects between both candidate updates tripled the instructions of a
// step.)
//
// The working type C is the factor's type: FP64 for the FP64 factorization,
// FP32 for the FP32 one (whose products elsewhere carry split-bf16 rounding
// far above FP32's, and which refinement in FP64 corrects).

// Waves of the diagonal-block kernel.  FP64: eight waves of eight column slots (two per SIMD).  FP32 (refined
// form): four waves of sixteen, one per SIMD at 124 registers, so that the
// workgroup fits beside a trailing-update wave of another frequency group
// (344 of a SIMD's 512 registers) instead of waiting for the update's
// workgroups to drain: 109-117 -> 92 us per launch beside the updates
// (44 -> 50 us alone).
template <typename T>
struct DiagShape {
  static constexpr int waves = 8;
};
template <>
struct DiagShape<float> {
  static constexpr int waves = 4;
};

template <typename C>
struct Zc {
  C re, im;
};
template <typename C>
__device__ __forceinline__ C fmac(C a, C b, C c) {
  if constexpr (std::is_same<C, float>::value) return __builtin_fmaf(a, b, c);
  else return __builtin_fma(a, b, c);
}
template <typename C>
__device__ __forceinline__ Zc<C> zmulc(Zc<C> a, Zc<C> b) {
  return {fmac(a.re, b.re, -(a.im * b.im)), fmac(a.re, b.im, a.im * b.re)};
}
template <typename C>
__device__ __forceinline__ Zc<C> zrecipc(Zc<C> d) {
  const C n = d.re * d.re + d.im * d.im;
  return {d.re / n, -d.im / n};
}

__device__ __forceinline__ double readlane_c(double v, int lane) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readlane(static_cast<int>(u & 0xffffffffu), lane);
  const uint32_t hi = __builtin_amdgcn_readlane(static_cast<int>(u >> 32), lane);
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}
__device__ __forceinline__ float readlane_c(float v, int lane) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}

// Slot update of the elimination step at column c: X role (column left of
// c: X[r][j] -= l X[c][j], X[c][j] from lane c) or A role (column right of
// c: A[r][j] -= l A[j][c], A[j][c] from the ring).
template <typename C>
__device__ __forceinline__ void diag_x(C &mr, C &mi, Zc<C> l, int c) {
  const Zc<C> pr = zmulc(l, Zc<C>{readlane_c(mr, c), readlane_c(mi, c)});
  mr = mr - pr.re;
  mi = mi - pr.im;
}
template <typename C>
__device__ __forceinline__ void diag_a(C &mr, C &mi, Zc<C> l, C br, C bi) {
  const Zc<C> pr = zmulc(l, Zc<C>{br, bi});
  mr = mr - pr.re;
  mi = mi - pr.im;
}

template <typename T>
__global__ void __launch_bounds__(64 * DiagShape<T>::waves) diag_kernel(T *__restrict__ re, T *__restrict__ im,
                                                   int64_t ld, int64_t mat_stride, int k0, int w,
                                                   int blk, int nblk, T *__restrict__ linvt_re,
                                                   T *__restrict__ linvt_im, T *__restrict__ dinv_re,
                                                   T *__restrict__ dinv_im,
                                                   int *__restrict__ singular,
                                                   const bf16x8 *__restrict__ lsplit,
                                                   const bf16x8 *__restrict__ wsplit,
                                                   int64_t split_stride, int G, int kc_beg,
                                                   int kc_end) {
  using C = T;
  constexpr int W = DiagShape<T>::waves, S = kB / W;
  const int f = blockIdx.x;
  T *Re = re + f * mat_stride, *Im = im + f * mat_stride;
  const int g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int r = threadIdx.x & 63;
  const int rc = min(r, w - 1);
  if constexpr (std::is_same<T, float>::value) {
    // left-looking path: the block's own update by the columns of its outer
    // block to its left (chunks [kc_beg, kc_end)), one 16-column strip per
    // wave of the first four, before the block is read
    if (kc_end > kc_beg) {
      if (g < 4)
        left_strip_update(Re, Im, ld, k0, k0 + w, k0 + 16 * g, k0 + w, kc_beg, kc_end,
                          lsplit + f * split_stride, wsplit + f * split_stride, split_stride, G, r);
      __syncthreads();
    }
  }
  __shared__ C cbr[2][kB], cbi[2][kB];
  // L11 by columns, D and D^-1: written to memory after the loop, so that
  // no global store is outstanding at the barriers of the column steps
  __shared__ C lsr[kB][kB], lsi[kB][kB], dsr[kB], dsi[kB], vsr[kB], vsi[kB];
  C mr[S], mi[S];
#pragma unroll
  for (int js = 0; js < S; ++js) {
    const int j = S * g + js;
    const bool in = r < w && j <= r;
    const int64_t a = static_cast<int64_t>(k0 + min(j, rc)) * ld + k0 + rc;
    const C vr = Re[a], vi = Im[a];
    mr[js] = in ? vr : C(0);
    mi[js] = in ? vi : C(0);
  }
  if (g == 0) {
    cbr[0][r] = mr[0];
    cbi[0][r] = mi[0];
  }
  const int64_t dbase = (static_cast<int64_t>(f) * nblk + blk) * kB;
  __syncthreads();
  for (int q = 0; q < W; ++q) {
#pragma unroll
    for (int t = 0; t < S; ++t) {
      const int c = S * q + t;
      if (c >= w) break;
      const int b = c & 1;
      const Zc<C> d{cbr[b][c], cbi[b][c]};
      const Zc<C> inv = zrecipc(d);
      const bool below = r > c && r < w;
      const Zc<C> lv = zmulc(Zc<C>{cbr[b][r], cbi[b][r]}, inv);
      const Zc<C> l{below ? lv.re : C(0), below ? lv.im : C(0)};
      if (g == 0) {
        lsr[c][r] = l.re;
        lsi[c][r] = l.im;
        if (r == 0) {
          dsr[c] = d.re;
          dsi[c] = d.im;
          vsr[c] = inv.re;
          vsi[c] = inv.im;
        }
      }
      if (g < q) {
#pragma unroll
        for (int js = 0; js < S; ++js) diag_x(mr[js], mi[js], l, c);
      } else if (g > q) {
        // all reads of the ring entry before its other entry is written
        C bcr[S], bci[S];
#pragma unroll
        for (int js = 0; js < S; ++js) {
          bcr[js] = cbr[b][S * g + js];
          bci[js] = cbi[b][S * g + js];
        }
#pragma unroll
        for (int js = 0; js < S; ++js) diag_a(mr[js], mi[js], l, bcr[js], bci[js]);
        if (t == S - 1 && g == q + 1) {  // column c + 1 = 8(q + 1): slot 0 here
          cbr[b ^ 1][r] = mr[0];
          cbi[b ^ 1][r] = mi[0];
        }
      } else {
        C bcr[S], bci[S];
#pragma unroll
        for (int js = t + 1; js < S; ++js) {
          bcr[js] = cbr[b][S * g + js];
          bci[js] = cbi[b][S * g + js];
        }
#pragma unroll
        for (int js = 0; js < S; ++js) {
          if (js < t) {
            diag_x(mr[js], mi[js], l, c);
          } else if (js == t) {
            mr[js] = (r == c ? C(1) : C(0)) - l.re;
            mi[js] = C(0) - l.im;
          } else {
            diag_a(mr[js], mi[js], l, bcr[js], bci[js]);
          }
        }
        if (t + 1 < S) {  // column c + 1 is this wave's next slot
          cbr[b ^ 1][r] = mr[t + 1];
          cbi[b ^ 1][r] = mi[t + 1];
        }
      }
      __syncthreads();
    }
  }
  const int64_t lbase = (static_cast<int64_t>(f) * nblk + blk) * kB * kB;
#pragma unroll
  for (int js = 0; js < S; ++js) {
    const int j = S * g + js;
    const bool in = r < w && j < w;
    linvt_re[lbase + j * kB + r] = in ? mr[js] : T(0);
    linvt_im[lbase + j * kB + r] = in ? mi[js] : T(0);
  }
  // L11 and D in place (column c, rows c..w-1), D^-1, the zero-pivot flag
  for (int q = threadIdx.x; q < kB * kB; q += blockDim.x) {
    const int c = q >> 6, rr = q & 63;
    if (c < w && rr >= c && rr < w) {
      const int64_t a = static_cast<int64_t>(k0 + c) * ld + k0 + rr;
      Re[a] = rr == c ? dsr[c] : lsr[c][rr];
      Im[a] = rr == c ? dsi[c] : lsi[c][rr];
    }
  }
  if (g == 0) {
    dinv_re[dbase + r] = r < w ? vsr[r] : T(0);
    dinv_im[dbase + r] = r < w ? vsi[r] : T(0);
    if (r < w && dsr[r] == C(0) && dsi[r] == C(0)) singular[f] = 1;
  }
}

// ---------------------------------------------------------------------------
// MFMA complex tile: D[i][j] = sum_k P[i][k] Q[k][j] over a 64 x 64 tile, i
// along the output's column axis and j along its row axis, so 16 consecutive
// output rows sit in 16 consecutive lanes (column-major stores are
// row-contiguous).  Staged operands in LDS: ps[k][i], qs[k][j].

template <typename T>
struct TileAcc {
  typename Mfma<T>::acc re[2][2], im[2][2];
};

template <typename T>
__device__ __forceinline__ void tile_zero(TileAcc<T> &a) {
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      a.re[x][y] = typename Mfma<T>::acc{0, 0, 0, 0};
      a.im[x][y] = typename Mfma<T>::acc{0, 0, 0, 0};
    }
}

template <typename T>
__device__ __forceinline__ void tile_mma(TileAcc<T> &a, const T (*pr)[kT], const T (*pi)[kT],
                                         const T (*qr)[kT], const T (*qi)[kT], int wi, int wj,
                                         int lane) {
  const int kq = lane >> 4, x = lane & 15;
#pragma unroll
  for (int kk = 0; kk < kKC; kk += 4) {
    T ar[2], ai[2], br[2], bi[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      ar[b] = pr[kk + kq][wi + 16 * b + x];
      ai[b] = pi[kk + kq][wi + 16 * b + x];
      br[b] = qr[kk + kq][wj + 16 * b + x];
      bi[b] = qi[kk + kq][wj + 16 * b + x];
    }
#pragma unroll
    for (int ib = 0; ib < 2; ++ib)
#pragma unroll
      for (int jb = 0; jb < 2; ++jb) {
        a.re[ib][jb] = Mfma<T>::run(ar[ib], br[jb], a.re[ib][jb]);
        a.re[ib][jb] = Mfma<T>::run(-ai[ib], bi[jb], a.re[ib][jb]);
        a.im[ib][jb] = Mfma<T>::run(ar[ib], bi[jb], a.im[ib][jb]);
        a.im[ib][jb] = Mfma<T>::run(ai[ib], br[jb], a.im[ib][jb]);
      }
  }
}

// Panel product of the refined form on the bf16 matrix cores, at the
// precision of the outer updates: W[r][c] = sum_k A21[r][k] Linv[c][k] over
// the block's k < w, each FP32 operand split as x = hi + lo and each real
// product formed as hi hi + hi lo + lo hi (twelve 16x16x32 instructions per
// complex 16 x 16 x 32 block).  Operands straight from memory into the
// operand order (A: lane l holds P[c = l % 16][k = 8 (l / 16) + e], B: lane l
// holds Q[k = 8 (l / 16) + e][r = l % 16]); P = Linv^T is zero outside the
// block, Q is masked to k < w and r < N.  The accumulators have the layout
// of the FP32 16x16x4 instruction's, so the stores that follow are shared.
__device__ __forceinline__ void panel_product_bf16(TileAcc<float> &acc, const float *__restrict__ lt_re,
                                                   const float *__restrict__ lt_im, int64_t lbase,
                                                   const float *__restrict__ Re,
                                                   const float *__restrict__ Im, int64_t ld, int k0,
                                                   int w, int r0, int N, int wi, int wj, int lane) {
  const int kq = lane >> 4, x16 = lane & 15;
#pragma unroll
  for (int ib = 0; ib < 2; ++ib)
#pragma unroll
    for (int jb = 0; jb < 2; ++jb) {
      acc.re[ib][jb] = f4{0, 0, 0, 0};
      acc.im[ib][jb] = f4{0, 0, 0, 0};
    }
#pragma unroll
  for (int kc = 0; kc < kB; kc += 32) {
    if (kc >= w) break;
    bf16x8 p[2][4], q[2][4];
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int k = kc + 8 * kq + e;
        const float ar = lt_re[lbase + k * kB + wi + 16 * b + x16];
        const float ai = lt_im[lbase + k * kB + wi + 16 * b + x16];
        const int rr = r0 + wj + 16 * b + x16;
        const int64_t kcl = min(k, w - 1), rcl = min(rr, N - 1);
        const float qa = Re[(k0 + kcl) * ld + rcl], qc = Im[(k0 + kcl) * ld + rcl];
        const bool ok = k < w && rr < N;
        __bf16 h, l;
        split_bf16(ar, h, l);
        p[b][0][e] = h;
        p[b][1][e] = l;
        split_bf16(ai, h, l);
        p[b][2][e] = h;
        p[b][3][e] = l;
        split_bf16(ok ? qa : 0.f, h, l);
        q[b][0][e] = h;
        q[b][1][e] = l;
        split_bf16(ok ? qc : 0.f, h, l);
        q[b][2][e] = h;
        q[b][3][e] = l;
      }
#pragma unroll
    for (int ib = 0; ib < 2; ++ib) {
      const bf16x8 nih = neg8(p[ib][2]), nil = neg8(p[ib][3]);
#pragma unroll
      for (int jb = 0; jb < 2; ++jb) {
        f4 cr = acc.re[ib][jb], ci = acc.im[ib][jb];
        cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[ib][0], q[jb][0], cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[ib][0], q[jb][1], cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[ib][1], q[jb][0], cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nih, q[jb][2], cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nih, q[jb][3], cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nil, q[jb][2], cr, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[ib][0], q[jb][2], ci, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[ib][0], q[jb][3], ci, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[ib][1], q[jb][2], ci, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[ib][2], q[jb][0], ci, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[ib][2], q[jb][1], ci, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[ib][3], q[jb][0], ci, 0, 0, 0);
        acc.re[ib][jb] = cr;
        acc.im[ib][jb] = ci;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// 2. Panel at k0 (width w, block blk): rows r in [k0 + w, N).
// W[r][c] = sum_k A21[r][k] Linv[c][k] (stored at W column wcol + c);
// L21[r][c] = W[r][c] Dinv[c] (stored in place).  D[i = c][j = r] with
// P[c][k] = linvt[k][c] and Q[k][r] = A21[r][k].

// At most 168 registers (three waves per SIMD), so that a panel workgroup
// fits beside a trailing-update wave of another frequency group (which
// holds 344 of a SIMD's 512 registers).
template <typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) panel_kernel(T *__restrict__ re, T *__restrict__ im,
                                                    int64_t ld, int64_t mat_stride, int k0, int w,
                                                    int blk, int nblk, int N, T *__restrict__ wre,
                                                    T *__restrict__ wim, int64_t w_stride, int wcol,
                                                    const T *__restrict__ linvt_re,
                                                    const T *__restrict__ linvt_im,
                                                    const T *__restrict__ dinv_re,
                                                    const T *__restrict__ dinv_im,
                                                    bf16x8 *__restrict__ lsplit,
                                                    bf16x8 *__restrict__ wsplit,
                                                    int64_t split_stride, int G, int write_w,
                                                    const bf16x8 *__restrict__ lsrc,
                                                    const bf16x8 *__restrict__ wsrc, int kc_beg,
                                                    int kc_end, int bf16_product, int ptiles,
                                                    int K) {
  // (frequency, row tile) units, frequency-major, dealt to the XCDs in
  // contiguous eighths: a frequency's tiles (sharing its L11^-1 and the
  // strip operands) meet in one L2, and the XCDs get equal shares
  const int units = ptiles * K, per = (units + 7) / 8;
  const int u = static_cast<int>(blockIdx.x % 8) * per + static_cast<int>(blockIdx.x / 8);
  if (u >= units) return;
  const int f = u / ptiles;
  T *Re = re + f * mat_stride, *Im = im + f * mat_stride;
  T *Wr = wre + f * w_stride, *Wi = wim + f * w_stride;
  const int64_t lbase = (static_cast<int64_t>(f) * nblk + blk) * kB * kB;
  const int64_t dbase = (static_cast<int64_t>(f) * nblk + blk) * kB;
  const int r0 = k0 + w + (u - f * ptiles) * kT;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if constexpr (std::is_same<T, float>::value) {
    // left-looking path: these panel rows' update by the columns of the
    // outer block left of this block (chunks [kc_beg, kc_end)), one
    // 16-column strip per wave, before the rows are read; the split copies
    // of earlier blocks, not the ones this launch writes
    if (kc_end > kc_beg) {
      left_strip_update(Re, Im, ld, r0, N, k0 + 16 * __builtin_amdgcn_readfirstlane(wave), k0 + w,
                        kc_beg, kc_end, lsrc + f * split_stride, wsrc + f * split_stride,
                        split_stride, G, lane);
      __syncthreads();
    }
  }
  const int wi = 32 * (wave >> 1), wj = 32 * (wave & 1);
  const int kq = lane >> 4, x16 = lane & 15;
  // Operands straight from memory into the matrix-core layout (lane l:
  // P[i = l % 16][k = l / 16], Q[k = l / 16][j = l % 16]), a 16-deep chunk
  // of k ahead of the products: no LDS staging and no barriers.  P is the
  // inverse (zero outside the block), Q the panel rows; the products run
  // in the order of the staged form, so W is unchanged.
  struct Chunk {
    T ar[4][2], ai[4][2], br[4][2], bi[4][2];
  };
  auto load = [&](int kc, Chunk &o) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = kc + 4 * q + kq;  // < kB: kc <= 48
      const int64_t kcl = min(k, w - 1);
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        o.ar[q][b] = linvt_re[lbase + k * kB + wi + 16 * b + x16];
        o.ai[q][b] = linvt_im[lbase + k * kB + wi + 16 * b + x16];
        const int rr = r0 + wj + 16 * b + x16;
        const int64_t rcl = min(rr, N - 1);
        const T a = Re[(k0 + kcl) * ld + rcl], c = Im[(k0 + kcl) * ld + rcl];
        const bool ok = k < w && rr < N;
        o.br[q][b] = ok ? a : T(0);
        o.bi[q][b] = ok ? c : T(0);
      }
    }
  };
  // D^-1 of this lane's output columns, loaded ahead of the products
  T dvr[2][4], dvi[2][4];
#pragma unroll
  for (int ib = 0; ib < 2; ++ib)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int c = wi + 16 * ib + Mfma<T>::row(lane, v);
      dvr[ib][v] = dinv_re[dbase + c];
      dvi[ib][v] = dinv_im[dbase + c];
    }
  TileAcc<T> acc;
  bool done = false;
  if constexpr (std::is_same<T, float>::value) {
    if (bf16_product) {
      panel_product_bf16(acc, linvt_re, linvt_im, lbase, Re, Im, ld, k0, w, r0, N, wi, wj, lane);
      done = true;
    }
  }
  if (!done) {
    tile_zero(acc);
    Chunk cur, nxt;
    load(0, cur);
#pragma unroll
    for (int c = 0; c < kB / 16; ++c) {
      const int kc = 16 * c;
      if (kc >= w) break;
      if (kc + 16 < w) load(kc + 16, nxt);
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int ib = 0; ib < 2; ++ib)
#pragma unroll
          for (int jb = 0; jb < 2; ++jb) {
            acc.re[ib][jb] = Mfma<T>::run(cur.ar[q][ib], cur.br[q][jb], acc.re[ib][jb]);
            acc.re[ib][jb] = Mfma<T>::run(-cur.ai[q][ib], cur.bi[q][jb], acc.re[ib][jb]);
            acc.im[ib][jb] = Mfma<T>::run(cur.ar[q][ib], cur.bi[q][jb], acc.im[ib][jb]);
            acc.im[ib][jb] = Mfma<T>::run(cur.ai[q][ib], cur.br[q][jb], acc.im[ib][jb]);
          }
      cur = nxt;
    }
  }
  // The results go in place: waves 0 and 2 (1 and 3) read the same 32 rows
  // over all 64 columns and each writes half of those columns.  Every
  // wave's loads come before any wave's stores, however far the waves have
  // drifted apart under other work on the compute unit.
  __syncthreads();
#pragma unroll
  for (int ib = 0; ib < 2; ++ib)
#pragma unroll
    for (int jb = 0; jb < 2; ++jb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int c = wi + 16 * ib + Mfma<T>::row(lane, v);
        const int64_t r = r0 + wj + 16 * jb + (lane & 15);
        if (c < w && r < N) {
          const T wr = acc.re[ib][jb][v], wim_ = acc.im[ib][jb][v];
          if (write_w) {
            Wr[(wcol + c) * ld + r] = wr;
            Wi[(wcol + c) * ld + r] = wim_;
          }
          const T dr = dvr[ib][v], di = dvi[ib][v];
          Re[(k0 + c) * ld + r] = wr * dr - wim_ * di;
          Im[(k0 + c) * ld + r] = wr * di + wim_ * dr;
        }
      }
  if constexpr (std::is_same<T, float>::value) {
    // Split copies of W and L for the left-looking update (w = 64): each
    // wave's 32 x 32 part (one 32-column chunk, two 16-row groups) goes
    // through LDS into the operand order, values as stored above.
    if (lsplit) {
      __shared__ float tsr[4][32][33], tsi[4][32][33];
      const int kc = (k0 + wi) / 32, kg = lane >> 4;
      for (int part = 0; part < 2; ++part) {
#pragma unroll
        for (int ib = 0; ib < 2; ++ib)
#pragma unroll
          for (int jb = 0; jb < 2; ++jb)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int cc = 16 * ib + Mfma<T>::row(lane, v), rr = 16 * jb + (lane & 15);
              const float wr = acc.re[ib][jb][v], wim_ = acc.im[ib][jb][v];
              float vr = wr, vi = wim_;
              if (part == 1) {
                const float dr = dvr[ib][v], di = dvi[ib][v];
                vr = wr * dr - wim_ * di;
                vi = wr * di + wim_ * dr;
              }
              tsr[wave][rr][cc] = vr;
              tsi[wave][rr][cc] = vi;
            }
        __syncthreads();
        bf16x8 *dst = (part == 0 ? wsplit : lsplit) + f * split_stride;
#pragma unroll
        for (int gg = 0; gg < 2; ++gg) {
          const int g = (r0 + wj) / 16 + gg;
          if (g < G) {
            const int rr = 16 * gg + (lane & 15);
            const bool rok = r0 + wj + rr < N;
            bf16x8 hr, lr, hi, li;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
              __bf16 h, l;
              split_bf16(rok ? tsr[wave][rr][8 * kg + e] : 0.f, h, l);
              hr[e] = h;
              lr[e] = l;
              split_bf16(rok ? tsi[wave][rr][8 * kg + e] : 0.f, h, l);
              hi[e] = h;
              li[e] = l;
            }
            const int64_t base = split_block(kc, g, G) * 256 + lane;
            dst[base] = hr;
            dst[base + 64] = lr;
            dst[base + 128] = hi;
            dst[base + 192] = li;
          }
        }
        __syncthreads();
      }
    }
  }
}

// ---------------------------------------------------------------------------
// 3. Update of the lower triangle in rows [rs, N) x columns [cs, ce):
// A[r][s] -= sum_{k < kn} W[r][wcol + k] L21[s][acol + k].
// D[i = s][j = r] with P[s][k] = L21[s][acol + k], Q[k][r] = W[r][wcol + k].
// triangular = 1 (cs == rs, ce == N): lower-triangular tile enumeration;
// otherwise a row-tile x column-tile grid whose tiles above the diagonal exit.

template <typename T>
__global__ void __launch_bounds__(256) update_kernel(T *__restrict__ re, T *__restrict__ im,
                                                     int64_t ld, int64_t mat_stride, int rs, int cs,
                                                     int ce, int N, int acol, int wcol, int kn,
                                                     const T *__restrict__ wre,
                                                     const T *__restrict__ wim, int64_t w_stride,
                                                     int triangular, int ctiles) {
  const int f = blockIdx.y;
  const int t = blockIdx.x;
  int tr, tc;
  if (triangular) {
    tr = static_cast<int>((__builtin_sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while (tr * (tr + 1) / 2 > t) --tr;
    while ((tr + 1) * (tr + 2) / 2 <= t) ++tr;
    tc = t - tr * (tr + 1) / 2;
  } else {
    tr = t / ctiles;
    tc = t % ctiles;
  }
  const int r0 = rs + tr * kT, c0 = cs + tc * kT;
  if (r0 + kT - 1 < c0) return;  // entirely above the diagonal
  T *Re = re + f * mat_stride, *Im = im + f * mat_stride;
  const T *Wr = wre + f * w_stride, *Wi = wim + f * w_stride;
  __shared__ T ps_r[kKC][kT], ps_i[kKC][kT], qs_r[kKC][kT], qs_i[kKC][kT];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wi = 32 * (wave >> 1), wj = 32 * (wave & 1);
  TileAcc<T> acc;
  tile_zero(acc);
  for (int kc = 0; kc < kn; kc += kKC) {
    __syncthreads();
    for (int q = threadIdx.x; q < kKC * kT; q += blockDim.x) {
      const int k = q / kT, x = q % kT;
      const int kk = kc + k;
      const bool ks = kk < kn, sok = c0 + x < ce, rok = r0 + x < N;
      const int64_t kcl = min(kk, kn - 1);
      const int64_t s = min(c0 + x, ce - 1), r = min(r0 + x, N - 1);
      const T a = Re[(acol + kcl) * ld + s], b = Im[(acol + kcl) * ld + s];
      const T c = Wr[(wcol + kcl) * ld + r], d = Wi[(wcol + kcl) * ld + r];
      ps_r[k][x] = (ks && sok) ? a : T(0);
      ps_i[k][x] = (ks && sok) ? b : T(0);
      qs_r[k][x] = (ks && rok) ? c : T(0);
      qs_i[k][x] = (ks && rok) ? d : T(0);
    }
    __syncthreads();
    tile_mma(acc, ps_r, ps_i, qs_r, qs_i, wi, wj, lane);
  }
  // C -= acc: all loads (clamped addresses) before the first subtraction.
  T cr[2][2][4], ci[2][2][4];
#pragma unroll
  for (int ib = 0; ib < 2; ++ib)
#pragma unroll
    for (int jb = 0; jb < 2; ++jb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t s = min(c0 + wi + 16 * ib + Mfma<T>::row(lane, v), ce - 1);
        const int64_t r = min(r0 + wj + 16 * jb + (lane & 15), N - 1);
        cr[ib][jb][v] = Re[s * ld + r];
        ci[ib][jb][v] = Im[s * ld + r];
      }
#pragma unroll
  for (int ib = 0; ib < 2; ++ib)
#pragma unroll
    for (int jb = 0; jb < 2; ++jb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t s = c0 + wi + 16 * ib + Mfma<T>::row(lane, v);
        const int64_t r = r0 + wj + 16 * jb + (lane & 15);
        if (s <= r && r < N && s < ce) {
          Re[s * ld + r] = cr[ib][jb][v] - acc.re[ib][jb][v];
          Im[s * ld + r] = ci[ib][jb][v] - acc.im[ib][jb][v];
        }
      }
}

// ---------------------------------------------------------------------------
// 3'. The outer update for FP32 factors on the bf16 matrix cores: each FP32
// operand x is split into bf16 parts x = hi + lo (+ a residual below
// 2^-16 |x|), and each real product P Q is formed as hi hi + hi lo + lo hi
// with FP32 accumulation (16x16x32 bf16 MFMA, 2.5 PFLOP/s dense on gfx950),
// twelve instructions per complex 16 x 16 x 32 block.  The dropped lo lo
// term and the split residuals bound the relative error of every product by
// about 3 * 2^-17; refinement in FP64 removes it.


// ---------------------------------------------------------------------------
// 3''. Left-looking form of the split-bf16 update.  Before the
// columns [J, J + nbj) are factorized they receive, in one launch, the
// contribution of the factorized columns k in [32 kc_beg, 32 kc_end):
//   A[r][s] -= sum_k W[r][k] L[s][k],   r >= s, s in [J, J + nbj),
// so every entry of A is read and written once by this kernel over the whole
// factorization instead of once per outer block.  The operands are the split
// copies below (written once per outer block), read straight into registers
// with one 16-byte load per lane and fragment: no LDS, no barriers.
//
// Split storage.  For each 32-column chunk kc of the factor and each 16-row
// group g >= 2 kc (the groups that meet the lower triangle of the chunk), a
// 4 KB block of four bf16 planes (re hi, re lo, im hi, im lo), each 64 lanes
// x 8 values in the operand order of the 16x16x32 bf16 MFMA: lane l holds
// row 16 g + l % 16, columns 32 kc + 8 (l / 16) + e.  Block index
// kc G - kc (kc - 1) + (g - 2 kc), G = ceil(N / 16).


// Tiles of the left-looking update: 128 x 128, rows r in [J, N), columns s
// in [J, J + nbj); tile (tr, tc) with tc < ctiles and tr >= tc, listed row
// by row, so the tiles of one row (sharing the W operand) are adjacent.  The
// launch's work units are the (frequency, tile) pairs, frequency-major, dealt
// to the XCDs in contiguous eighths (the hardware assigns workgroup b to XCD
// b % 8): every XCD gets the same number of units to within one, and the
// tiles of one frequency (sharing its L operand) meet in one XCD's L2.
// Dealt by tile alone, a launch whose tile count is 8 k - 5 left one XCD five
// tiles short for every frequency, idle for part of the launch.
__device__ __forceinline__ bool left_unit(int b, int rtiles, int ctiles, int K, int &f, int &tr,
                                          int &tc) {
  const int tri = ctiles * (ctiles - 1) / 2;
  const int T = tri + (rtiles - (ctiles - 1)) * ctiles;
  const int units = T * K;
  const int per = (units + 7) / 8;
  const int u = (b % 8) * per + b / 8;
  if (u >= units) return false;
  f = u / T;
  const int t = u - f * T;
  if (t < tri) {
    tr = static_cast<int>((__builtin_sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while (tr * (tr + 1) / 2 > t) --tr;
    while ((tr + 1) * (tr + 2) / 2 <= t) ++tr;
    tc = t - tr * (tr + 1) / 2;
  } else {
    const int v = t - tri;
    tr = ctiles - 1 + v / ctiles;
    tc = v % ctiles;
  }
  return true;
}

__global__ void __launch_bounds__(256) update_left_kernel(
    float *__restrict__ re, float *__restrict__ im, int64_t ld, int64_t mat_stride, int J,
    int nbj, int N, int G, int rtiles, int ctiles, int kc_beg, int kc_end,
    const bf16x8 *__restrict__ lsplit, const bf16x8 *__restrict__ wsplit, int64_t split_stride,
    int narrow, int K, const double *__restrict__ lmesh, int64_t loops,
    const double *__restrict__ omegas, const int64_t *__restrict__ rlc_ptr,
    const int32_t *__restrict__ rlc_row, const double *__restrict__ rlc_val, int r_band,
    int m_int) {
  int f, tr, tc;
  if (!left_unit(blockIdx.x, rtiles, ctiles, K, f, tr, tc)) return;
  // The wave index is uniform across the wave: keeping it (and every
  // address derived from it) in scalar registers leaves the vector
  // registers to the operands and accumulators.
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  // 128 x 128 tiles (2 x 2 waves), or for a block of at most 64 columns
  // 256 x 64 tiles (the four waves down the rows, sharing the L operand)
  const int wi = narrow ? 0 : 64 * (wave >> 1), wj = narrow ? 64 * wave : 64 * (wave & 1);
  const int s0 = J + 128 * tc + wi, r0 = J + (narrow ? 256 : 128) * tr + wj;
  // Operands through buffer loads: one descriptor per operand and
  // frequency (scalar), a fixed per-lane offset, the block offset in a
  // scalar register -- no per-lane 64-bit address arithmetic.
  const __amdgpu_buffer_rsrc_t rL = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<bf16x8 *>(lsplit + f * split_stride), 0, static_cast<int>(split_stride * 16),
      0x00020000);
  const __amdgpu_buffer_rsrc_t rW = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<bf16x8 *>(wsplit + f * split_stride), 0, static_cast<int>(split_stride * 16),
      0x00020000);
  const int voff = lane * 16;
  int gs[4], gr[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    gs[b] = min(s0 / 16 + b, G - 1);
    gr[b] = min(r0 / 16 + b, G - 1);
  }
  f4 accr[4][4], acci[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      accr[a][b] = f4{0, 0, 0, 0};
      acci[a][b] = f4{0, 0, 0, 0};
    }
  bf16x8 p[4][4], qa[4][4], qb[4][4];
  auto load_p = [&](int kc, int b) {
    const int so = static_cast<int>(split_block(kc, gs[b], G) * 4096);
#pragma unroll
    for (int pl = 0; pl < 4; ++pl)
      p[b][pl] = __builtin_bit_cast(
          bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rL, voff + 1024 * pl, so, 0));
  };
  auto load_q = [&](int kc, bf16x8 (&q)[4][4]) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int so = static_cast<int>(split_block(kc, gr[b], G) * 4096);
#pragma unroll
      for (int pl = 0; pl < 4; ++pl)
        q[b][pl] = __builtin_bit_cast(
            bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rW, voff + 1024 * pl, so, 0));
    }
  };
  const int nkc = kc_end;
  // One chunk of k: the 16-row group ib of the L operand is refilled with
  // chunk kc + 1 as soon as its products are issued, so the loads of the
  // next chunk are in flight behind the remaining matrix-core work (the W
  // operand is double-buffered; loads past the last chunk re-read it).
  auto step = [&](int kc, const bf16x8 (&q)[4][4]) {
    const int kn = min(kc + 1, nkc - 1);
#pragma unroll
    for (int ib = 0; ib < 4; ++ib) {
      const bf16x8 nih = neg8(p[ib][2]), nil = neg8(p[ib][3]);
#pragma unroll
      for (int jb = 0; jb < 4; ++jb) {
        f4 cr = accr[ib][jb], ci = acci[ib][jb];
        cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[ib][0], q[jb][0], cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[ib][0], q[jb][1], cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[ib][1], q[jb][0], cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nih, q[jb][2], cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nih, q[jb][3], cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nil, q[jb][2], cr, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[ib][0], q[jb][2], ci, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[ib][0], q[jb][3], ci, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[ib][1], q[jb][2], ci, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[ib][2], q[jb][0], ci, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[ib][2], q[jb][1], ci, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[ib][3], q[jb][0], ci, 0, 0, 0);
        accr[ib][jb] = cr;
        acci[ib][jb] = ci;
      }
      load_p(kn, ib);
      // keep the refill here, ahead of the next group's products
      __builtin_amdgcn_sched_barrier(0);
    }
  };
#pragma unroll
  for (int b = 0; b < 4; ++b) load_p(kc_beg, b);
  load_q(kc_beg, qa);
  // kc_end - kc_beg is even: the loop body has no branch for the compiler
  // to sink the refills into.
  for (int kc = kc_beg; kc < nkc; kc += 2) {
    load_q(kc + 1, qb);
    __builtin_amdgcn_sched_barrier(0);
    step(kc, qa);
    load_q(min(kc + 2, nkc - 1), qa);
    __builtin_amdgcn_sched_barrier(0);
    step(kc + 1, qb);
  }
  const int kq = lane >> 4, x16 = lane & 15;
  const int se = J + nbj;
#pragma unroll
  for (int ib = 0; ib < 4; ++ib) {
    float cr[4][4], ci[4][4];
    if (lmesh) {
      // This launch is the only one that touches these entries: it forms
      // R + j w L itself, the values the form and scatter kernels would
      // have stored (j w L rounded to FP32, the FP32 of the r_mesh entry
      // or 0), instead of reading them back from the matrix.
      const double w = omegas[f];
#pragma unroll
      for (int jb = 0; jb < 4; ++jb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int64_t s = min(s0 + 16 * ib + 4 * kq + v, N - 1);
          const int64_t r = min(r0 + 16 * jb + x16, N - 1);
          cr[jb][v] = 0.f;
          ci[jb][v] = static_cast<float>(w * lmesh[s * loops + r]);
        }
      // r_mesh entries lie within r_band below the diagonal in the interior
      // rows, anywhere in the port rows: most of the wave's 64 x 16
      // sub-blocks hold none
      const bool any_r = r0 - (s0 + 16 * ib + 15) <= r_band || r0 + 64 > m_int;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int s = s0 + 16 * ib + 4 * kq + v;
        if (!any_r || s >= N) continue;
        for (int64_t e = rlc_ptr[s]; e < rlc_ptr[s + 1]; ++e) {
          const int d = rlc_row[e] - r0 - x16;  // 16 jb for this lane's rows
          if (d >= 0 && d < 64 && (d & 15) == 0) {
            const float rv = static_cast<float>(rlc_val[e]);
#pragma unroll
            for (int jb = 0; jb < 4; ++jb)
              if (d == 16 * jb) cr[jb][v] = rv;
          }
        }
      }
    } else {
#pragma unroll
      for (int jb = 0; jb < 4; ++jb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int64_t s = min(s0 + 16 * ib + 4 * kq + v, N - 1);
          const int64_t r = min(r0 + 16 * jb + x16, N - 1);
          cr[jb][v] = re[f * mat_stride + s * ld + r];
          ci[jb][v] = im[f * mat_stride + s * ld + r];
        }
    }
#pragma unroll
    for (int jb = 0; jb < 4; ++jb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t s = s0 + 16 * ib + 4 * kq + v;
        const int64_t r = r0 + 16 * jb + x16;
        if (s <= r && r < N && s < se) {
          re[f * mat_stride + s * ld + r] = cr[jb][v] - accr[ib][jb][v];
          im[f * mat_stride + s * ld + r] = ci[jb][v] - acci[ib][jb][v];
        }
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

// ---------------------------------------------------------------------------
// Refinement.  Vectors are FP64, [2 v][m] for re and [2 v + 1][m] for im,
// v = f p + q (frequency f, port column q).  Solves with the stored factors
// go by super-blocks of 256 rows, two launches each over all vectors: the
// diagonal part (one workgroup per vector; four 64-row sub-blocks, each a
// product with the stored inverse of its unit-lower block and an update of
// the sub-blocks after it), then the off-diagonal update of the other rows
// (forward: the rows below, one row per thread, the block's columns read
// along rows; backward: the rows above, through 64 x 64 tiles of L staged in
// LDS so that the transposed product reads L along its columns).
constexpr int kSB = 256;

// Forward, diagonal part of the super-block at B0 (bw rows): solves
// L_BB z = y_B for y_B = work[B0, B0 + bw) and writes z to out.
template <typename T>
__global__ void __launch_bounds__(256) fwd_diag_kernel(
    const T *__restrict__ re, const T *__restrict__ im, int64_t ld, int64_t mat_stride, int m,
    int p, int B0, int bw, int nblk, const T *__restrict__ linvt_re,
    const T *__restrict__ linvt_im, const double *__restrict__ work, double *__restrict__ out) {
  const int v = blockIdx.x, f = v / p;
  const T *Re = re + f * mat_stride, *Im = im + f * mat_stride;
  const double *yr = work + static_cast<int64_t>(2 * v) * m, *yi = yr + m;
  double *orr = out + static_cast<int64_t>(2 * v) * m, *oi = orr + m;
  __shared__ double ysr[kSB], ysi[kSB], zsr[kB], zsi[kB], ptr_[4][kB], pti[4][kB];
  const int tid = threadIdx.x;
  ysr[tid] = tid < bw ? yr[B0 + tid] : 0.0;
  ysi[tid] = tid < bw ? yi[B0 + tid] : 0.0;
  __syncthreads();
  const int i = tid & 63, q = tid >> 6;
  for (int s0 = 0; s0 < bw; s0 += kB) {
    const int sw = min(kB, bw - s0), blk = (B0 + s0) / kB;
    const int64_t lbase = (static_cast<int64_t>(f) * nblk + blk) * kB * kB;
    // z_i = sum_j Linv[i][j] y_j, Linv[i][j] = linvt[j][i] (unit lower,
    // zero outside the block)
    double ar = 0.0, ai = 0.0;
#pragma unroll 4
    for (int jj = 0; jj < 16; ++jj) {
      const int j = 16 * q + jj;
      const double lr = static_cast<double>(linvt_re[lbase + j * kB + i]);
      const double li = static_cast<double>(linvt_im[lbase + j * kB + i]);
      const double y_r = ysr[s0 + j], y_i = ysi[s0 + j];
      ar += lr * y_r - li * y_i;
      ai += lr * y_i + li * y_r;
    }
    ptr_[q][i] = ar;
    pti[q][i] = ai;
    __syncthreads();
    if (tid < kB) {
      const double zr = (ptr_[0][tid] + ptr_[1][tid]) + (ptr_[2][tid] + ptr_[3][tid]);
      const double zi = (pti[0][tid] + pti[1][tid]) + (pti[2][tid] + pti[3][tid]);
      zsr[tid] = zr;
      zsi[tid] = zi;
      if (tid < sw) {
        orr[B0 + s0 + tid] = zr;
        oi[B0 + s0 + tid] = zi;
      }
    }
    __syncthreads();
    // later rows of the super-block: y[r] -= sum_c L[r][s0 + c] z[c]
    const int r = s0 + sw + tid;
    if (r < bw) {
      double ur = ysr[r], ui = ysi[r];
      for (int c = 0; c < sw; ++c) {
        const int64_t a = static_cast<int64_t>(B0 + s0 + c) * ld + B0 + r;
        const double lr = static_cast<double>(Re[a]), li = static_cast<double>(Im[a]);
        ur -= lr * zsr[c] - li * zsi[c];
        ui -= lr * zsi[c] + li * zsr[c];
      }
      ysr[r] = ur;
      ysi[r] = ui;
    }
    __syncthreads();
  }
}

// Forward, rows below the super-block: work[r] -= sum_c L[r][B0 + c] out[B0 + c].
// 64 rows per workgroup, four threads per row (a quarter of the columns
// each, loads issued in batches of 16), partial sums added in fixed order.
template <typename T>
__global__ void __launch_bounds__(256) fwd_update_kernel(
    const T *__restrict__ re, const T *__restrict__ im, int64_t ld, int64_t mat_stride, int m,
    int p, int B0, int bw, double *__restrict__ work, const double *__restrict__ out) {
  const int v = blockIdx.y, f = v / p;
  const T *Re = re + f * mat_stride, *Im = im + f * mat_stride;
  double *yr = work + static_cast<int64_t>(2 * v) * m, *yi = yr + m;
  const double *orr = out + static_cast<int64_t>(2 * v) * m, *oi = orr + m;
  __shared__ double zr[kSB], zi[kSB], ptr_[4][kB], pti[4][kB];
  const int tid = threadIdx.x, lane = tid & 63, q = tid >> 6;
  zr[tid] = tid < bw ? orr[B0 + tid] : 0.0;  // zero past bw
  zi[tid] = tid < bw ? oi[B0 + tid] : 0.0;
  __syncthreads();
  const int r = B0 + bw + blockIdx.x * kB + lane;
  const int64_t rc = min(r, m - 1);
  double ar = 0.0, ai = 0.0;
#pragma unroll
  for (int c0 = 0; c0 < kB; c0 += 16) {
    T lr[16], li[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int64_t c = min(kB * q + c0 + e, bw - 1);
      lr[e] = Re[(B0 + c) * ld + rc];
      li[e] = Im[(B0 + c) * ld + rc];
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int c = kB * q + c0 + e;
      const double a = static_cast<double>(lr[e]), b = static_cast<double>(li[e]);
      ar += a * zr[c] - b * zi[c];
      ai += a * zi[c] + b * zr[c];
    }
  }
  ptr_[q][lane] = ar;
  pti[q][lane] = ai;
  __syncthreads();
  if (q == 0 && r < m) {
    yr[r] -= (ptr_[0][lane] + ptr_[1][lane]) + (ptr_[2][lane] + ptr_[3][lane]);
    yi[r] -= (pti[0][lane] + pti[1][lane]) + (pti[2][lane] + pti[3][lane]);
  }
}

// z = y D^-1 (D on the factor's diagonal), from `in` into `out`.
template <typename T>
__global__ void dscale_kernel(const T *__restrict__ re, const T *__restrict__ im, int64_t ld,
                              int64_t mat_stride, int m, int p, int V, const double *__restrict__ in,
                              double *__restrict__ out) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<int64_t>(V) * m) return;
  const int v = static_cast<int>(t / m), i = static_cast<int>(t % m);
  const int f = v / p;
  const int64_t a = f * mat_stride + static_cast<int64_t>(i) * ld + i;
  const Z2 inv = zrecip(Z2{static_cast<double>(re[a]), static_cast<double>(im[a])});
  const int64_t o = static_cast<int64_t>(2 * v) * m + i;
  const Z2 y = zmul(Z2{in[o], in[o + m]}, inv);
  out[o] = y.re;
  out[o + m] = y.im;
}

// Backward, diagonal part of the super-block at B0: solves L_BB^T x = z_B
// for z_B = work[B0, B0 + bw) (sub-blocks last to first) and writes x to out.
template <typename T>
__global__ void __launch_bounds__(256) bwd_diag_kernel(
    const T *__restrict__ re, const T *__restrict__ im, int64_t ld, int64_t mat_stride, int m,
    int p, int B0, int bw, int nblk, const T *__restrict__ linvt_re,
    const T *__restrict__ linvt_im, const double *__restrict__ work, double *__restrict__ out) {
  const int v = blockIdx.x, f = v / p;
  const T *Re = re + f * mat_stride, *Im = im + f * mat_stride;
  const double *zr_g = work + static_cast<int64_t>(2 * v) * m, *zi_g = zr_g + m;
  double *orr = out + static_cast<int64_t>(2 * v) * m, *oi = orr + m;
  __shared__ double zsr[kSB], zsi[kSB], xsr[kB], xsi[kB], ptr_[4][kB], pti[4][kB];
  const int tid = threadIdx.x;
  zsr[tid] = tid < bw ? zr_g[B0 + tid] : 0.0;
  zsi[tid] = tid < bw ? zi_g[B0 + tid] : 0.0;
  __syncthreads();
  const int c = tid & 63, q = tid >> 6;
  for (int s0 = ((bw - 1) / kB) * kB; s0 >= 0; s0 -= kB) {
    const int sw = min(kB, bw - s0), blk = (B0 + s0) / kB;
    const int64_t lbase = (static_cast<int64_t>(f) * nblk + blk) * kB * kB;
    // x_c = sum_i Linv[i][c] z_i, Linv[i][c] = linvt[c][i]
    double ar = 0.0, ai = 0.0;
#pragma unroll 4
    for (int ii = 0; ii < 16; ++ii) {
      const int i = 16 * q + ii;
      const double lr = static_cast<double>(linvt_re[lbase + c * kB + i]);
      const double li = static_cast<double>(linvt_im[lbase + c * kB + i]);
      const double z_r = zsr[s0 + i], z_i = zsi[s0 + i];
      ar += lr * z_r - li * z_i;
      ai += lr * z_i + li * z_r;
    }
    ptr_[q][c] = ar;
    pti[q][c] = ai;
    __syncthreads();
    if (tid < kB) {
      const double xr = (ptr_[0][tid] + ptr_[1][tid]) + (ptr_[2][tid] + ptr_[3][tid]);
      const double xi = (pti[0][tid] + pti[1][tid]) + (pti[2][tid] + pti[3][tid]);
      xsr[tid] = xr;
      xsi[tid] = xi;
      if (tid < sw) {
        orr[B0 + s0 + tid] = xr;
        oi[B0 + s0 + tid] = xi;
      }
    }
    __syncthreads();
    // earlier rows of the super-block: z[i] -= sum_c L[s0 + c][i] x[c]
    if (tid < s0) {
      const int i = tid;
      double ur = zsr[i], ui = zsi[i];
      const int64_t base = static_cast<int64_t>(B0 + i) * ld + B0 + s0;
#pragma unroll 8
      for (int cc = 0; cc < sw; ++cc) {
        const double lr = static_cast<double>(Re[base + cc]), li = static_cast<double>(Im[base + cc]);
        ur -= lr * xsr[cc] - li * xsi[cc];
        ui -= lr * xsi[cc] + li * xsr[cc];
      }
      zsr[i] = ur;
      zsi[i] = ui;
    }
    __syncthreads();
  }
}

// Backward, rows above the super-block: work[i] -= sum_c L[B0 + c][i] out[B0 + c]
// for i < B0.  Workgroup of 64 rows i; L[B0 + c][i] = A(row B0 + c, col i)
// is staged as 64 x 64 tiles, read along c (contiguous), used along i.
template <typename T>
__global__ void __launch_bounds__(256) bwd_update_kernel(
    const T *__restrict__ re, const T *__restrict__ im, int64_t ld, int64_t mat_stride, int m,
    int p, int B0, int bw, double *__restrict__ work, const double *__restrict__ out) {
  const int v = blockIdx.y, f = v / p;
  const T *Re = re + f * mat_stride, *Im = im + f * mat_stride;
  double *zr_g = work + static_cast<int64_t>(2 * v) * m, *zi_g = zr_g + m;
  const double *orr = out + static_cast<int64_t>(2 * v) * m, *oi = orr + m;
  __shared__ double xr[kSB], xi[kSB], ptr_[4][kB], pti[4][kB];
  __shared__ T tr_[kB][kB + 1], ti_[kB][kB + 1];  // [i][c]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  xr[tid] = tid < bw ? orr[B0 + tid] : 0.0;
  xi[tid] = tid < bw ? oi[B0 + tid] : 0.0;
  const int i0 = blockIdx.x * kB;
  double ar = 0.0, ai = 0.0;
  // 16-byte loads: V consecutive rows B0 + c of column i per lane, 64 / V
  // lanes per column, V * 64 / 64 columns i per wave instruction; entries
  // past bw are zeros (the rows below the super-block are not read)
  constexpr int V = 16 / static_cast<int>(sizeof(T));
  typedef T TV __attribute__((ext_vector_type(V)));
  constexpr int LPC = kB / V;  // lanes per column i
  const int cq_last = ((bw - 1) / V) * V;
  for (int c0 = 0; c0 < bw; c0 += kB) {
    __syncthreads();
    const int cl = V * (lane % LPC);
    const int cq = c0 + cl;
#pragma unroll
    for (int k = 0; k < 16 * LPC / 64; ++k) {
      const int ii = 16 * wave + (64 / LPC) * k + lane / LPC;
      const int i = min(i0 + ii, B0 - 1);
      const int64_t a = static_cast<int64_t>(i) * ld + B0 + min(cq, cq_last);
      const TV vr = *reinterpret_cast<const TV *>(Re + a), vi = *reinterpret_cast<const TV *>(Im + a);
#pragma unroll
      for (int e = 0; e < V; ++e) {
        const bool ok = cq + e < bw;
        tr_[ii][cl + e] = ok ? vr[e] : T(0);
        ti_[ii][cl + e] = ok ? vi[e] : T(0);
      }
    }
    __syncthreads();
    // thread (i = lane, quarter = wave): 16 columns of this tile
#pragma unroll 4
    for (int k = 0; k < 16; ++k) {
      const int cc = 16 * wave + k;
      const double lr = static_cast<double>(tr_[lane][cc]), li = static_cast<double>(ti_[lane][cc]);
      const double x_r = xr[c0 + cc], x_i = xi[c0 + cc];  // zero past bw
      ar += lr * x_r - li * x_i;
      ai += lr * x_i + li * x_r;
    }
  }
  ptr_[wave][lane] = ar;
  pti[wave][lane] = ai;
  __syncthreads();
  const int i = i0 + tid;
  if (tid < kB && i < B0) {
    zr_g[i] -= (ptr_[0][tid] + ptr_[1][tid]) + (ptr_[2][tid] + ptr_[3][tid]);
    zi_g[i] -= (pti[0][tid] + pti[1][tid]) + (pti[2][tid] + pti[3][tid]);
  }
}

// P[q][i] = sum_j L[j][i] X[q][j], i < m, over Q = 2 K p vectors; L is the
// symmetric FP64 l_mesh (leading dimension lds), read column-contiguously.
// The j range is cut into `slices` (so the launch fills the chip); slice s
// writes its partial sums to part[s][q][i] and lx_reduce_kernel adds the
// slices in order.
__global__ void __launch_bounds__(256) lx_kernel(const double *__restrict__ L, int64_t lds, int m,
                                                 int Q, const double *__restrict__ X,
                                                 double *__restrict__ part, int slice_len) {
  const int i0 = blockIdx.x * 64, q0 = blockIdx.y * 64, sl = blockIdx.z;
  const int jb = sl * slice_len, je = min(m, jb + slice_len);
  __shared__ double ls[32][65], xs[32][65];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  double acc[4][4] = {};
  for (int j0 = jb; j0 < je; j0 += 32) {
    __syncthreads();
    for (int q = threadIdx.x; q < 32 * 64; q += blockDim.x) {
      const int j = q / 64, x = q % 64;
      const int jj = j0 + j;
      const bool lok = jj < je && i0 + x < m;
      const double lv = L[static_cast<int64_t>(min(jj, m - 1)) * lds + min(i0 + x, m - 1)];
      ls[j][x] = lok ? lv : 0.0;
      const int xq = q / 32, xj = q % 32;
      const bool xok = j0 + xj < je && q0 + xq < Q;
      const double xv = X[static_cast<int64_t>(min(q0 + xq, Q - 1)) * m + min(j0 + xj, m - 1)];
      xs[xj][xq] = xok ? xv : 0.0;
    }
    __syncthreads();
    for (int j = 0; j < 32; ++j) {
      double a[4], b[4];
      for (int u = 0; u < 4; ++u) a[u] = ls[j][tx + 16 * u];
      for (int u = 0; u < 4; ++u) b[u] = xs[j][ty + 16 * u];
      for (int u = 0; u < 4; ++u)
        for (int s = 0; s < 4; ++s) acc[u][s] = __builtin_fma(a[u], b[s], acc[u][s]);
    }
  }
  double *out = part + static_cast<int64_t>(sl) * Q * m;
  for (int u = 0; u < 4; ++u)
    for (int s = 0; s < 4; ++s) {
      const int i = i0 + tx + 16 * u, q = q0 + ty + 16 * s;
      if (i < m && q < Q) out[static_cast<int64_t>(q) * m + i] = acc[u][s];
    }
}

__global__ void lx_reduce_kernel(int64_t n, int slices, const double *__restrict__ part,
                                 double *__restrict__ P) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  double a = part[t];
  for (int s = 1; s < slices; ++s) a += part[s * n + t];
  P[t] = a;
}

// r = b - (R + j w L) x with b = Zep: r_re = b_re - (R x)_re + w P_im,
// r_im = w l_ep - (R x)_im - w P_re.
__global__ void residual_kernel(int m, int p, int V, const double *__restrict__ omegas,
                                const int64_t *__restrict__ rptr, const int32_t *__restrict__ rcol,
                                const double *__restrict__ rval, const double *__restrict__ rep,
                                const double *__restrict__ lmesh, int64_t lds,
                                const double *__restrict__ X, const double *__restrict__ P,
                                double *__restrict__ Rv) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<int64_t>(V) * m) return;
  const int v = static_cast<int>(t / m), i = static_cast<int>(t % m);
  const int f = v / p, q = v % p;
  const double w = omegas[f];
  const double *xr = X + static_cast<int64_t>(2 * v) * m, *xi = xr + m;
  double sr = 0.0, si = 0.0;
  for (int64_t e = rptr[i]; e < rptr[i + 1]; ++e) {
    sr += rval[e] * xr[rcol[e]];
    si += rval[e] * xi[rcol[e]];
  }
  const double b_re = rep[static_cast<int64_t>(q) * m + i];
  const double b_im = w * lmesh[static_cast<int64_t>(m + q) * lds + i];
  const double pr = P[static_cast<int64_t>(2 * v) * m + i], pim = P[static_cast<int64_t>(2 * v + 1) * m + i];
  Rv[static_cast<int64_t>(2 * v) * m + i] = b_re - sr + w * pim;
  Rv[static_cast<int64_t>(2 * v + 1) * m + i] = b_im - si - w * pr;
}

// R = b for X = 0: what residual_kernel gives there (L X = +0, so
// R_re = b_re - 0 + w 0 = b_re + 0 and R_im = b_im - 0 - w 0 = b_im), without
// the product L X.
__global__ void border_rhs_kernel(int m, int p, int V, const double *__restrict__ omegas,
                                  const double *__restrict__ rep, const double *__restrict__ lmesh,
                                  int64_t lds, double *__restrict__ Rv) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<int64_t>(V) * m) return;
  const int v = static_cast<int>(t / m), i = static_cast<int>(t % m);
  const int f = v / p, q = v % p;
  const double w = omegas[f];
  Rv[static_cast<int64_t>(2 * v) * m + i] = rep[static_cast<int64_t>(q) * m + i] + 0.0;
  Rv[static_cast<int64_t>(2 * v + 1) * m + i] = w * lmesh[static_cast<int64_t>(m + q) * lds + i];
}

__global__ void axpy_copy_kernel(int64_t n, double *__restrict__ x, const double *__restrict__ dx,
                                 double *__restrict__ next, const double *__restrict__ r) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  if (dx) x[t] += dx[t];
  if (next) next[t] = r[t];
}

// x += dx for the vectors of the frequencies still refining (active[f]):
// the steps after the first ones run for a whole group of frequencies, and a
// frequency already accepted keeps its x (and Z) whatever its neighbours in
// the group need, so its result does not depend on how the sweep is grouped.
__global__ void active_axpy_kernel(int m, int p, int64_t n, double *__restrict__ x,
                                   const double *__restrict__ dx, const int *__restrict__ active) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  const int f = static_cast<int>(t / (2 * static_cast<int64_t>(m))) / p;
  if (active[f]) x[t] += dx[t];
}

// Z[f][q1][q2] = Zpp - sum_i b^{q1}_i x^{q2}_i - sum_i x^{q1}_i r^{q2}_i and
// the relative size of the last (first-order) term; one block per entry,
// fixed-order reduction.
__global__ void __launch_bounds__(256) functional_kernel(
    int m, int p, const double *__restrict__ omegas, const double *__restrict__ rep,
    const double *__restrict__ rpp, const double *__restrict__ lmesh, int64_t lds,
    const double *__restrict__ X, const double *__restrict__ Rv, double *__restrict__ z,
    double *__restrict__ metric, const int *__restrict__ active) {
  const int e = blockIdx.x;
  const int f = e / (p * p), q1 = (e / p) % p, q2 = e % p;
  // (a frequency already accepted keeps its Z; the whole block leaves)
  if (active && !active[f]) return;
  const double w = omegas[f];
  const int v1 = f * p + q1, v2 = f * p + q2;
  const double *x1r = X + static_cast<int64_t>(2 * v1) * m, *x1i = x1r + m;
  const double *x2r = X + static_cast<int64_t>(2 * v2) * m, *x2i = x2r + m;
  const double *r2r = Rv + static_cast<int64_t>(2 * v2) * m, *r2i = r2r + m;
  double a[4] = {0, 0, 0, 0};  // b.x re, im; x.r re, im
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    const double bre = rep[static_cast<int64_t>(q1) * m + i];
    const double bim = w * lmesh[static_cast<int64_t>(m + q1) * lds + i];
    a[0] += bre * x2r[i] - bim * x2i[i];
    a[1] += bre * x2i[i] + bim * x2r[i];
    a[2] += x1r[i] * r2r[i] - x1i[i] * r2i[i];
    a[3] += x1r[i] * r2i[i] + x1i[i] * r2r[i];
  }
  __shared__ double red[4][256];
  for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = a[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s)
      for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double zr = rpp[q1 * p + q2] - red[0][0] - red[2][0];
    const double zi = w * lmesh[static_cast<int64_t>(m + q1) * lds + m + q2] - red[1][0] - red[3][0];
    z[2 * (static_cast<int64_t>(f) * p * p + q1 * p + q2)] = zr;
    z[2 * (static_cast<int64_t>(f) * p * p + q1 * p + q2) + 1] = w == 0.0 ? 0.0 : zi;
    const double corr = __builtin_sqrt(red[2][0] * red[2][0] + red[3][0] * red[3][0]);
    const double zz = __builtin_sqrt(zr * zr + zi * zi);
    metric[e] = zz > 0.0 ? corr / zz : corr;
  }
}

// Squared norms of the last correction dx and of the residual r of every
// vector (re and im rows adjacent): nrm[2 v] = |dx_v|^2, nrm[2 v + 1] =
// |r_v|^2; fixed-order reduction.
__global__ void __launch_bounds__(256) norms_kernel(int m, const double *__restrict__ DX,
                                                    const double *__restrict__ Rv,
                                                    double *__restrict__ nrm) {
  const int v = blockIdx.x;
  const double *d = DX + static_cast<int64_t>(2 * v) * m, *r = Rv + static_cast<int64_t>(2 * v) * m;
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < 2 * m; i += blockDim.x) {
    a += d[i] * d[i];
    b += r[i] * r[i];
  }
  __shared__ double ra[256], rb[256];
  ra[threadIdx.x] = a;
  rb[threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s) {
      ra[threadIdx.x] += ra[threadIdx.x + s];
      rb[threadIdx.x] += rb[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    nrm[2 * v] = ra[0];
    nrm[2 * v + 1] = rb[0];
  }
}

// ---------------------------------------------------------------------------
// GMRES-IR (Carson & Higham 2017-2018): the correction equation Zee d = r of
// each refinement step solved by GMRES on M^-1 Zee d = M^-1 r, M the FP32
// factorization, everything but M in FP64.  The V complex vectors of a batch
// run in lockstep, each with its own Krylov basis and Hessenberg system (on
// the host); vector v's real row is 2 v, its imaginary row 2 v + 1, as in X.

// Y = (R + j w L) X given P = L X.
__global__ void zee_apply_kernel(int m, int p, int V, const double *__restrict__ omegas,
                                 const int64_t *__restrict__ rptr, const int32_t *__restrict__ rcol,
                                 const double *__restrict__ rval, const double *__restrict__ X,
                                 const double *__restrict__ P, double *__restrict__ Y) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<int64_t>(V) * m) return;
  const int v = static_cast<int>(t / m), i = static_cast<int>(t % m);
  const double w = omegas[v / p];
  const double *xr = X + static_cast<int64_t>(2 * v) * m, *xi = xr + m;
  double sr = 0.0, si = 0.0;
  for (int64_t e = rptr[i]; e < rptr[i + 1]; ++e) {
    sr += rval[e] * xr[rcol[e]];
    si += rval[e] * xi[rcol[e]];
  }
  const double pr = P[static_cast<int64_t>(2 * v) * m + i], pim = P[static_cast<int64_t>(2 * v + 1) * m + i];
  Y[static_cast<int64_t>(2 * v) * m + i] = sr - w * pim;
  Y[static_cast<int64_t>(2 * v + 1) * m + i] = si + w * pr;
}

// h[v][i] = <b_i, w>_v = sum conj(b_i) w over vector v, for i < nb; one
// block per (i, v), fixed-order reduction.  Basis vector i starts at
// basis + i * bstride.
__global__ void __launch_bounds__(256) cdot_kernel(int m, int hld, const double *__restrict__ basis,
                                                   int64_t bstride, const double *__restrict__ W,
                                                   double *__restrict__ h) {
  const int i = blockIdx.x, v = blockIdx.y;
  const double *br = basis + i * bstride + static_cast<int64_t>(2 * v) * m, *bi = br + m;
  const double *wr = W + static_cast<int64_t>(2 * v) * m, *wi = wr + m;
  double a = 0.0, b = 0.0;
  for (int t = threadIdx.x; t < m; t += blockDim.x) {
    a += br[t] * wr[t] + bi[t] * wi[t];
    b += br[t] * wi[t] - bi[t] * wr[t];
  }
  __shared__ double ra[256], rb[256];
  ra[threadIdx.x] = a;
  rb[threadIdx.x] = b;
  __syncthreads();
  for (int s2 = 128; s2 > 0; s2 >>= 1) {
    if (static_cast<int>(threadIdx.x) < s2) {
      ra[threadIdx.x] += ra[threadIdx.x + s2];
      rb[threadIdx.x] += rb[threadIdx.x + s2];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    h[2 * (static_cast<int64_t>(v) * hld + i)] = ra[0];
    h[2 * (static_cast<int64_t>(v) * hld + i) + 1] = rb[0];
  }
}

// W_v -= sum_{i < nb} c[v][i] b_i (sign = -1), or D_v = sum_{i < nb} c[v][i] b_i
// (sign = +1, W ignored on input), c complex with leading dimension cld.
__global__ void basis_combine_kernel(int m, int V, int nb, const double *__restrict__ basis,
                                     int64_t bstride, const double *__restrict__ c, int cld,
                                     double *__restrict__ W, int sign) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<int64_t>(V) * m) return;
  const int v = static_cast<int>(t / m), k = static_cast<int>(t % m);
  const int64_t o = static_cast<int64_t>(2 * v) * m + k;
  double ar = 0.0, ai = 0.0;
  for (int i = 0; i < nb; ++i) {
    const double hr = c[2 * (static_cast<int64_t>(v) * cld + i)];
    const double hi = c[2 * (static_cast<int64_t>(v) * cld + i) + 1];
    const double br = basis[i * bstride + o], bi = basis[i * bstride + o + m];
    ar += hr * br - hi * bi;
    ai += hr * bi + hi * br;
  }
  if (sign < 0) {
    W[o] -= ar;
    W[o + m] -= ai;
  } else {
    W[o] = ar;
    W[o + m] = ai;
  }
}

// nrm[v] = |W_v|^2, fixed-order reduction.
__global__ void __launch_bounds__(256) vnorm_kernel(int m, const double *__restrict__ W,
                                                    double *__restrict__ nrm) {
  const int v = blockIdx.x;
  const double *a = W + static_cast<int64_t>(2 * v) * m;
  double s2 = 0.0;
  for (int t = threadIdx.x; t < 2 * m; t += blockDim.x) s2 += a[t] * a[t];
  __shared__ double r[256];
  r[threadIdx.x] = s2;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (static_cast<int>(threadIdx.x) < k) r[threadIdx.x] += r[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) nrm[v] = r[0];
}

// dst_v = W_v / sqrt(nrm[v]) (0 where the norm is 0 or not finite).
__global__ void vscale_kernel(int m, int V, const double *__restrict__ W,
                              const double *__restrict__ nrm, double *__restrict__ dst) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<int64_t>(2) * V * m) return;
  const int v = static_cast<int>(t / (2 * static_cast<int64_t>(m)));
  const double q = nrm[v];
  const double inv = q > 0.0 && q < INFINITY ? 1.0 / __builtin_sqrt(q) : 0.0;
  dst[t] = W[t] * inv;
}

// Zeroes the vectors of the frequencies flagged in `reset`.
__global__ void reset_vectors_kernel(int m, int p, int V, const int *__restrict__ reset,
                                     double *__restrict__ X) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<int64_t>(2) * V * m) return;
  const int v = static_cast<int>(t / (2 * static_cast<int64_t>(m)));
  if (reset[v / p]) X[t] = 0.0;
}

// ---------------------------------------------------------------------------
// Device-side state of the captured refinement steps: per frequency f,
// st[0 K + f] converged (0/1), st[1 K + f] steps, st[2 K + f] the last
// first-order term |x^T r| / |Z|, st[3 K + f] its value one step earlier,
// st[4 K + f] |dx| one step earlier, st[5 K + f] the first |dx|.

__global__ void fill_kernel(int64_t n, double *__restrict__ x, double v) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t < n) x[t] = v;
}

__global__ void fill_int_kernel(int n, int *__restrict__ x, int v) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) x[t] = v;
}

__global__ void refine_state_init_kernel(int K, double *__restrict__ st) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= K) return;
  st[f] = 0.0;
  st[K + f] = 0.0;
  st[2 * K + f] = INFINITY;
  st[3 * K + f] = INFINITY;
  st[4 * K + f] = 0.0;
  st[5 * K + f] = 0.0;
}

// The acceptance test of step `it` of the classical refinement, the host's
// test of ldlt_refined_batch operation for operation (same sums in the same
// order), for the frequencies not yet converged.
__global__ void refine_accept_kernel(int K, int p, int it, const double *__restrict__ metric,
                                     const double *__restrict__ nrm, const double *__restrict__ z,
                                     const int *__restrict__ sing, double z_tol, double tol,
                                     double *__restrict__ st) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= K || st[f] != 0.0) return;
  double worst = 0.0;
  for (int q = 0; q < p * p; ++q) worst = fmax(worst, metric[f * p * p + q]);
  if (!(worst == worst)) worst = INFINITY;
  st[K + f] += 1.0;
  st[2 * K + f] = worst;
  double sdx = 0.0, sr = 0.0, sz = 0.0;
  for (int q = 0; q < p; ++q) {
    sdx += nrm[2 * (f * p + q)];
    sr += nrm[2 * (f * p + q) + 1];
  }
  for (int q = 0; q < 2 * p * p; ++q) {
    const double zq = z[2 * f * p * p + q];
    sz += zq * zq;
  }
  const double ndx = sqrt(sdx), nr = sqrt(sr), nz = sqrt(sz);
  double bound = INFINITY;
  if (st[5 * K + f] == 0.0 && isfinite(ndx)) st[5 * K + f] = ndx;
  const double pdx = st[4 * K + f];
  if (it >= 2 && pdx > 0.0) {
    const double rho = ndx / pdx;
    const double x_bound = rho <= 0.5 ? rho / (1.0 - rho) * ndx : INFINITY;
    if (rho <= 0.5 && nz > 0.0) bound = x_bound * nr / nz;
  }
  st[4 * K + f] = ndx;
  const double prev = st[3 * K + f];
  const bool done = isfinite(nz) && sing[f] == 0 &&
                    (bound <= z_tol || worst <= tol || (worst <= 1e-9 && !(worst < 0.5 * prev)));
  if (done) st[f] = 1.0;
  st[3 * K + f] = worst;
}

// ---------------------------------------------------------------------------
// Host drivers.

struct Layout {
  int m, p, N, nblk, nb;
  int64_t ld, mat_stride, w_stride;
};

Layout layout_of(const DenseSolveInputs &in, int nb) {
  Layout L;
  L.m = static_cast<int>(in.m);
  L.p = static_cast<int>(in.p);
  L.N = L.m + L.p;
  L.nb = nb;
  L.nblk = (L.m + kB - 1) / kB;
  L.ld = ((static_cast<int64_t>(L.N) + 63) / 64) * 64;
  L.mat_stride = L.ld * L.ld;
  L.w_stride = L.ld * nb;
  return L;
}

template <typename T>
struct Factor {
  T *re, *im, *wre, *wim, *lt_re, *lt_im, *di_re, *di_im;
};

template <typename T>
size_t factor_bytes(const Layout &L, int K) {
  const size_t per = static_cast<size_t>(L.mat_stride) * 2 + static_cast<size_t>(L.w_stride) * 2 +
                     static_cast<size_t>(std::max(L.nblk, 1)) * (2 * kB * kB + 2 * kB);
  return static_cast<size_t>(K) * per * sizeof(T);
}

template <typename T>
Factor<T> carve(char *base, const Layout &L, int K) {
  Factor<T> F;
  T *p = reinterpret_cast<T *>(base);
  const size_t nb = static_cast<size_t>(std::max(L.nblk, 1));
  F.re = p;
  p += static_cast<size_t>(K) * L.mat_stride;
  F.im = p;
  p += static_cast<size_t>(K) * L.mat_stride;
  F.wre = p;
  p += static_cast<size_t>(K) * L.w_stride;
  F.wim = p;
  p += static_cast<size_t>(K) * L.w_stride;
  F.lt_re = p;
  p += static_cast<size_t>(K) * nb * kB * kB;
  F.lt_im = p;
  p += static_cast<size_t>(K) * nb * kB * kB;
  F.di_re = p;
  p += static_cast<size_t>(K) * nb * kB;
  F.di_im = p;
  return F;
}

template <typename T>
void launch_update(const Layout &L, const Factor<T> &F, int K, int rs, int cs, int ce, int acol,
                   int wcol, int kn, bool triangular, hipStream_t stream, Profiler *prof) {
  const int rtiles = (L.N - rs + kT - 1) / kT;
  const int ctiles = (ce - cs + kT - 1) / kT;
  if (rtiles <= 0 || ctiles <= 0 || kn <= 0) return;
  const int tiles = triangular ? rtiles * (rtiles + 1) / 2 : rtiles * ctiles;
  if (prof) prof->begin(stream);
  update_kernel<T><<<dim3(tiles, K), 256, 0, stream>>>(F.re, F.im, L.ld, L.mat_stride, rs, cs, ce,
                                                       L.N, acol, wcol, kn, F.wre, F.wim,
                                                       L.w_stride, triangular ? 1 : 0, ctiles);
  if (prof) {
    // algorithmic flops: 8 per complex multiply-add over the lower-triangle
    // entries of the region
    double entries = 0.0;
    for (int s = cs; s < ce; ++s) entries += static_cast<double>(L.N - std::max(s, rs));
    prof->end(stream, 8.0 * kn * entries * K);
  }
}

// Split operand copies of the left-looking form, per frequency.
struct SplitBufs {
  bf16x8 *l = nullptr, *w = nullptr;
  int G = 0;
  int64_t stride = 0;  // bf16x8 elements per frequency
};

int64_t split_blocks(const Layout &L) {
  const int G = (L.N + 15) / 16, KC = (L.m + 31) / 32;
  return split_block(KC, 2 * KC, G);
}

// With `form` (its r_mesh columns and l_mesh) the launch also forms the
// entries it updates, which no earlier kernel has touched.
fh_status launch_left_update(const Layout &L, const Factor<float> &F, const SplitBufs &sb, int K,
                             int J, int nbj, int kc_beg, int kc_end, hipStream_t stream,
                             Profiler *prof, const DenseSolveInputs *form,
                             const double *d_omegas) {
  // the kernel takes k in pairs of 32-column chunks
  if (kc_end <= kc_beg || (kc_end - kc_beg) % 2 != 0)
    return set_error(FH_E_INVALID_ARGUMENT, "left-looking update: odd chunk range");
  // a block of at most 64 columns: 256-row tiles one block wide
  const int narrow = nbj <= 64 ? 1 : 0;
  const int rtiles = narrow ? (L.N - J + 255) / 256 : (L.N - J + 127) / 128;
  const int ctiles = narrow ? 1 : (nbj + 127) / 128;
  const int T = ctiles * (ctiles - 1) / 2 + (rtiles - (ctiles - 1)) * ctiles;
  const int64_t units = static_cast<int64_t>(T) * K;
  if (units > 0x7fffff00)
    return set_error(FH_E_INVALID_ARGUMENT, "left-looking update: too many tiles for one launch");
  if (prof) prof->begin(stream);
  update_left_kernel<<<static_cast<unsigned>((units + 7) / 8 * 8), 256, 0, stream>>>(
      F.re, F.im, L.ld, L.mat_stride, J, nbj, L.N, sb.G, rtiles, ctiles, kc_beg, kc_end, sb.l,
      sb.w, sb.stride, narrow, K, form ? form->d_lmesh : nullptr, static_cast<int64_t>(L.N),
      d_omegas, form ? form->d_rlc_ptr : nullptr, form ? form->d_rlc_row : nullptr,
      form ? form->d_rlc_val : nullptr, form ? form->r_band : 0, L.m);
  if (prof) {
    double entries = 0.0;
    for (int s = J; s < J + nbj; ++s) entries += static_cast<double>(L.N - s);
    prof->end(stream, 8.0 * 32.0 * (kc_end - kc_beg) * entries * K, 3.0);
  }
  FH_HIP_TRY(hipGetLastError());
  return FH_OK;
}

// Factorizes (L D L^T) the first m pivots of the bordered matrices.  With
// `sb` (FP32 only) the outer updates are left-looking on the split-bf16
// copies; the trailing p x p block is then not updated (the refined form
// takes Z from the refinement, not from the factor).
template <typename T>
fh_status factorize(const DenseSolveInputs &in, const Layout &L, const double *d_omegas, int K,
                    const Factor<T> &F, int *d_singular, hipStream_t stream, Profiler *prof,
                    const SplitBufs *sb = nullptr) {
  // On the left-looking path every column right of the first outer block
  // is first touched by its trailing update, which forms it; only the first
  // block is formed here.
  const bool update_forms = std::is_same<T, float>::value && sb && in.d_rlc_ptr && L.m > L.nb;
  const int cend = update_forms ? L.nb : L.N;
  {
    // chunks of 16 bytes: a grid row holds at most N / V + 4 of them; the
    // columns c < cend are the grid rows y = c, or y = N - 1 - c when
    // cend > (N + 1) / 2
    constexpr int V = 16 / static_cast<int>(sizeof(T));
    dim3 grid(((L.N + V - 1) / V + 4 + 255) / 256, std::min((L.N + 1) / 2, cend));
    form_kernel<T><<<grid, 256, 0, stream>>>(in.d_lmesh, static_cast<int64_t>(L.N), d_omegas, K,
                                             L.N, L.ld, F.re, F.im, L.mat_stride, cend);
    FH_HIP_TRY(hipGetLastError());
    if (in.r_nnz) {
      const size_t blocks = std::min<size_t>((in.r_nnz + 255) / 256, 4096);
      dim3 g2(static_cast<unsigned>(blocks), K);
      scatter_r_kernel<T><<<g2, 256, 0, stream>>>(in.d_r_rows, in.d_r_cols, in.d_r_vals,
                                                  static_cast<int64_t>(in.r_nnz), L.ld, F.re,
                                                  L.mat_stride, cend);
      FH_HIP_TRY(hipGetLastError());
    }
  }
  for (int K0 = 0; K0 < L.m; K0 += L.nb) {
    const int KE = std::min(K0 + L.nb, L.m);
    if constexpr (std::is_same<T, float>::value) {
      if (sb && K0 > 0) {
        const fh_status st = launch_left_update(L, F, *sb, K, K0, KE - K0, 0, K0 / 32, stream, prof,
                                                update_forms ? &in : nullptr, d_omegas);
        if (st) return st;
      }
    }
    for (int k0 = K0; k0 < KE; k0 += kB) {
      const int w = std::min(kB, KE - k0);
      const int blk = k0 / kB;
      // left-looking inside the outer block too: the earlier 64-column
      // blocks of this outer block, from their split copies, applied by the
      // diagonal-block and panel kernels to their own rows
      const int ikc_beg = sb && k0 > K0 ? K0 / 32 : 0, ikc_end = sb && k0 > K0 ? k0 / 32 : 0;
      diag_kernel<T><<<K, 64 * DiagShape<T>::waves, 0, stream>>>(F.re, F.im, L.ld, L.mat_stride, k0, w, blk, L.nblk,
                                            F.lt_re, F.lt_im, F.di_re, F.di_im, d_singular,
                                            sb ? sb->l : nullptr, sb ? sb->w : nullptr,
                                            sb ? sb->stride : 0, sb ? sb->G : 0, ikc_beg, ikc_end);
      FH_HIP_TRY(hipGetLastError());
      const int rows = L.N - (k0 + w);
      if (rows <= 0) continue;
      const int ptiles = (rows + kT - 1) / kT;
      // on the left-looking path the panel also writes the split copies of
      // its W and L (needed when columns remain), and no FP32 W
      const bool emit = sb && k0 + w < L.m;
      const unsigned pblocks = static_cast<unsigned>((static_cast<int64_t>(ptiles) * K + 7) / 8 * 8);
      panel_kernel<T><<<pblocks, 256, 0, stream>>>(
          F.re, F.im, L.ld, L.mat_stride, k0, w, blk, L.nblk, L.N, F.wre, F.wim, L.w_stride,
          k0 - K0, F.lt_re, F.lt_im, F.di_re, F.di_im, emit ? sb->l : nullptr,
          emit ? sb->w : nullptr, sb ? sb->stride : 0, sb ? sb->G : 0, sb ? 0 : 1,
          sb ? sb->l : nullptr, sb ? sb->w : nullptr, ikc_beg, ikc_end, sb ? 1 : 0, ptiles, K);
      FH_HIP_TRY(hipGetLastError());
      if (sb) continue;
      if (k0 + w < KE)
        launch_update<T>(L, F, K, k0 + w, k0 + w, KE, k0, k0 - K0, w, false, stream, nullptr);
      FH_HIP_TRY(hipGetLastError());
    }
    if (KE < L.N) {
      if constexpr (std::is_same<T, float>::value) {
        if (sb) continue;
      }
      launch_update<T>(L, F, K, KE, KE, L.N, K0, 0, KE - K0, true, stream, prof);
    }
    FH_HIP_TRY(hipGetLastError());
  }
  return FH_OK;
}

// The backward half of a solve: vec <- L^-T work (work: the forward
// substitution already scaled by D^-1).
template <typename T>
fh_status factor_solve_bwd(const Layout &L, const Factor<T> &F, int V, double *vec, double *work,
                           hipStream_t stream) {
  const int m = L.m;
  const int nsb = (m + kSB - 1) / kSB;
  for (int sb = nsb - 1; sb >= 0; --sb) {
    const int B0 = sb * kSB, bw = std::min(kSB, m - B0);
    bwd_diag_kernel<T><<<V, 256, 0, stream>>>(F.re, F.im, L.ld, L.mat_stride, m, L.p, B0, bw,
                                              L.nblk, F.lt_re, F.lt_im, work, vec);
    if (B0 > 0)
      bwd_update_kernel<T><<<dim3((B0 + kB - 1) / kB, V), 256, 0, stream>>>(
          F.re, F.im, L.ld, L.mat_stride, m, L.p, B0, bw, work, vec);
  }
  FH_HIP_TRY(hipGetLastError());
  return FH_OK;
}

// v <- Zee^-1 v for all V vectors with the FP32 factors; `work` is scratch
// of the same size.
template <typename T>
fh_status factor_solve(const Layout &L, const Factor<T> &F, int V, double *vec, double *work,
                       hipStream_t stream) {
  const int m = L.m;
  const size_t n = static_cast<size_t>(2) * V * m;
  axpy_copy_kernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, stream>>>(
      static_cast<int64_t>(n), nullptr, nullptr, work, vec);
  const int nsb = (m + kSB - 1) / kSB;
  for (int sb = 0; sb < nsb; ++sb) {
    const int B0 = sb * kSB, bw = std::min(kSB, m - B0);
    fwd_diag_kernel<T><<<V, 256, 0, stream>>>(F.re, F.im, L.ld, L.mat_stride, m, L.p, B0, bw,
                                              L.nblk, F.lt_re, F.lt_im, work, vec);
    const int below = m - B0 - bw;
    if (below > 0)
      fwd_update_kernel<T><<<dim3((below + kB - 1) / kB, V), 256, 0, stream>>>(
          F.re, F.im, L.ld, L.mat_stride, m, L.p, B0, bw, work, vec);
  }
  FH_HIP_TRY(hipGetLastError());
  dscale_kernel<T><<<static_cast<unsigned>((static_cast<size_t>(V) * m + 255) / 256), 256, 0, stream>>>(
      F.re, F.im, L.ld, L.mat_stride, m, L.p, V, vec, work);
  FH_HIP_TRY(hipGetLastError());
  return factor_solve_bwd<T>(L, F, V, vec, work, stream);
}


}  // namespace

int ldlt_default_nb(bool refined) { return refined ? kNbRefined : kNbPlain; }

size_t ldlt_bytes_per_frequency(size_t m, size_t p, bool refined, int nb) {
  DenseSolveInputs in;
  in.m = m;
  in.p = p;
  if (!refined) return factor_bytes<double>(layout_of(in, nb > 0 ? nb : kNbPlain), 1);
  const Layout L = layout_of(in, nb > 0 ? nb : kNbRefined);
  // the FP64 vectors of the refinement: X, R, DX, P, W, up to 16 partial
  // sums of L X and the GMRES-IR basis, 2 p rows of m each
  return factor_bytes<float>(L, 1) +
         static_cast<size_t>(p) * m * 2 * (5 + 16 + kGmresRestart + 1 + 2) * sizeof(double) +
         2 * static_cast<size_t>(split_blocks(L)) * 4096;
}

fh_status ldlt_solve_batch(const DenseSolveInputs &in, const double *d_omegas, int K, double *d_z,
                           BatchResult &out, LdltWorkspace &ws, hipStream_t stream,
                           Profiler *prof) {
  const Layout L = layout_of(in, in.nb > 0 ? in.nb : kNbPlain);
  fh_status st;
  const size_t fbytes = factor_bytes<double>(L, K);
  if ((st = ws.reserve(fbytes + static_cast<size_t>(K) * sizeof(int)))) return st;
  const Factor<double> F = carve<double>(ws.base, L, K);
  int *d_sing = reinterpret_cast<int *>(ws.base + fbytes);
  FH_HIP_TRY(hipMemsetAsync(d_sing, 0, K * sizeof(int), stream));
  if ((st = factorize<double>(in, L, d_omegas, K, F, d_sing, stream, prof))) return st;
  extract_kernel<double><<<K, 64, 0, stream>>>(F.re, F.im, L.ld, L.mat_stride, L.m, L.p, d_omegas,
                                               d_z);
  FH_HIP_TRY(hipGetLastError());
  out.singular.assign(K, 0);
  FH_HIP_TRY(hipMemcpyAsync(out.singular.data(), d_sing, K * sizeof(int), hipMemcpyDeviceToHost,
                            stream));
  FH_HIP_TRY(hipStreamSynchronize(stream));
  out.refine_steps.assign(K, 0);
  out.refine_change.assign(K, 0.0);
  out.converged.assign(K, 1);
  out.stage.assign(K, -1);
  return FH_OK;
}

fh_status ldlt_refined_batch(const DenseSolveInputs &in, const double *d_omegas, int K,
                             double *d_z, BatchResult &out, LdltWorkspace &ws,
                             hipStream_t stream, Profiler *prof) {
  const Layout L = layout_of(in, in.nb > 0 ? in.nb : kNbRefined);
  const int m = L.m, p = L.p, V = K * p, Q = 2 * V;
  fh_status st;
  const size_t fbytes = factor_bytes<float>(L, K);
  const size_t vbytes = static_cast<size_t>(Q) * m * sizeof(double);
  const int itiles = (m + 63) / 64, qtiles = (Q + 63) / 64;
  const int lx_slices = std::max(1, std::min(16, 2048 / std::max(1, itiles * qtiles)));
  const int lx_slice_len = ((m + lx_slices - 1) / lx_slices + 31) / 32 * 32;
  // X, R, DX, P, the lx slices, the GMRES basis and its work vector, then
  // the small arrays
  const int hld = kGmresRestart + 1;
  const size_t small = static_cast<size_t>(K) * p * p * sizeof(double) +  // metric
                       2 * static_cast<size_t>(V) * sizeof(double) +     // nrm
                       static_cast<size_t>(V) * sizeof(double) +         // vnrm
                       4 * static_cast<size_t>(V) * hld * sizeof(double) +  // h (2 passes)
                       2 * static_cast<size_t>(V) * hld * sizeof(double) +  // y
                       6 * static_cast<size_t>(K) * sizeof(double) +        // device refinement state
                       3 * static_cast<size_t>(K) * sizeof(int) + 256;      // sing, reset, active
  const size_t extra = (5 + static_cast<size_t>(lx_slices) + hld) * vbytes + small;
  const size_t sbytes = static_cast<size_t>(split_blocks(L)) * 4096 * K;
  if ((st = ws.reserve(fbytes + extra + 2 * sbytes + 512))) return st;
  const Factor<float> F = carve<float>(ws.base, L, K);
  SplitBufs sb;
  sb.G = (L.N + 15) / 16;
  sb.stride = split_blocks(L) * 256;
  {
    char *sp = ws.base + ((fbytes + extra + 255) / 256) * 256;
    sb.l = reinterpret_cast<bf16x8 *>(sp);
    sb.w = reinterpret_cast<bf16x8 *>(sp + sbytes);
  }
  const size_t n = static_cast<size_t>(Q) * m;
  char *cur = ws.base + ((fbytes + 255) / 256) * 256;
  double *X = reinterpret_cast<double *>(cur);
  double *R = X + n;
  double *DX = R + n;
  double *P = DX + n;
  double *Wv = P + n;
  double *lx_part = Wv + n;
  double *basis = lx_part + static_cast<size_t>(lx_slices) * n;
  double *metric = basis + static_cast<size_t>(hld) * n;
  double *d_nrm = metric + static_cast<size_t>(K) * p * p;
  double *d_vnrm = d_nrm + 2 * static_cast<size_t>(V);
  double *d_h = d_vnrm + V;
  double *d_y = d_h + 4 * static_cast<size_t>(V) * hld;
  double *d_state = d_y + 2 * static_cast<size_t>(V) * hld;
  int *d_sing = reinterpret_cast<int *>(d_state + 6 * static_cast<size_t>(K));
  int *d_reset = d_sing + K;
  int *d_active = d_reset + K;
  const unsigned vec_blocks = static_cast<unsigned>((n + 255) / 256);
  const unsigned res_blocks = static_cast<unsigned>((static_cast<size_t>(V) * m + 255) / 256);
  const int64_t lds = L.N;
  std::vector<double> h_metric(static_cast<size_t>(K) * p * p), h_nrm(2 * static_cast<size_t>(V));
  std::vector<double> h_z(2 * static_cast<size_t>(K) * p * p);
  std::vector<int> h_sing(K, 0), h_reset(K, 0), h_active(K, 1);

  // P = L X, then R = b - Zee X.
  auto residual = [&]() -> fh_status {
    lx_kernel<<<dim3(itiles, qtiles, lx_slices), 256, 0, stream>>>(in.d_lmesh, lds, m, Q, X, lx_part,
                                                                   lx_slice_len);
    lx_reduce_kernel<<<vec_blocks, 256, 0, stream>>>(static_cast<int64_t>(n), lx_slices, lx_part, P);
    residual_kernel<<<res_blocks, 256, 0, stream>>>(m, p, V, d_omegas, in.d_ree_ptr, in.d_ree_col,
                                                    in.d_ree_val, in.d_rep, in.d_lmesh, lds, X, P, R);
    FH_HIP_TRY(hipGetLastError());
    return FH_OK;
  };

  // DX ~= Zee^-1 R by GMRES on M^-1 Zee d = M^-1 R from d = 0: one cycle of
  // up to kGmresRestart steps, stopped per vector when the preconditioned
  // residual falls below kGmresTol of its start.  P and Wv are scratch.
  auto gmres_correction = [&](int &steps_out) -> fh_status {
    typedef std::complex<double> cd;
    const int R1 = kGmresRestart;
    std::vector<double> beta(V), hh(4 * static_cast<size_t>(V) * hld), vn(V);
    std::vector<cd> H(static_cast<size_t>(V) * (R1 + 1) * R1), g(static_cast<size_t>(V) * (R1 + 1));
    std::vector<double> cs(static_cast<size_t>(V) * R1);
    std::vector<cd> sn(static_cast<size_t>(V) * R1);
    std::vector<int> done(V, 0), used(V, 0);
    // b_0 = M^-1 R / beta
    FH_HIP_TRY(hipMemcpyAsync(Wv, R, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
    fh_status s2 = factor_solve<float>(L, F, V, Wv, P, stream);
    if (s2) return s2;
    vnorm_kernel<<<V, 256, 0, stream>>>(m, Wv, d_vnrm);
    vscale_kernel<<<vec_blocks, 256, 0, stream>>>(m, V, Wv, d_vnrm, basis);
    FH_HIP_TRY(hipGetLastError());
    FH_HIP_TRY(hipMemcpyAsync(vn.data(), d_vnrm, V * sizeof(double), hipMemcpyDeviceToHost, stream));
    FH_HIP_TRY(hipStreamSynchronize(stream));
    for (int v = 0; v < V; ++v) {
      beta[v] = std::sqrt(vn[v]);
      g[static_cast<size_t>(v) * (R1 + 1)] = cd(beta[v], 0.0);
      if (!(beta[v] > 0.0) || !std::isfinite(beta[v])) done[v] = 1;
    }
    int j = 0;
    for (; j < R1; ++j) {
      bool all = true;
      for (int v = 0; v < V; ++v) all = all && done[v];
      if (all) break;
      const double *bj = basis + static_cast<size_t>(j) * n;
      // W = M^-1 Zee b_j
      lx_kernel<<<dim3(itiles, qtiles, lx_slices), 256, 0, stream>>>(in.d_lmesh, lds, m, Q, bj, lx_part,
                                                                     lx_slice_len);
      lx_reduce_kernel<<<vec_blocks, 256, 0, stream>>>(static_cast<int64_t>(n), lx_slices, lx_part, P);
      zee_apply_kernel<<<res_blocks, 256, 0, stream>>>(m, p, V, d_omegas, in.d_ree_ptr, in.d_ree_col,
                                                       in.d_ree_val, bj, P, Wv);
      FH_HIP_TRY(hipGetLastError());
      if ((s2 = factor_solve<float>(L, F, V, Wv, P, stream))) return s2;
      // classical Gram-Schmidt, twice
      for (int pass = 0; pass < 2; ++pass) {
        double *hp = d_h + static_cast<size_t>(pass) * 2 * V * hld;
        cdot_kernel<<<dim3(j + 1, V), 256, 0, stream>>>(m, hld, basis, static_cast<int64_t>(n), Wv, hp);
        basis_combine_kernel<<<res_blocks, 256, 0, stream>>>(m, V, j + 1, basis, static_cast<int64_t>(n),
                                                             hp, hld, Wv, -1);
      }
      vnorm_kernel<<<V, 256, 0, stream>>>(m, Wv, d_vnrm);
      vscale_kernel<<<vec_blocks, 256, 0, stream>>>(m, V, Wv, d_vnrm,
                                                    basis + static_cast<size_t>(j + 1) * n);
      FH_HIP_TRY(hipGetLastError());
      FH_HIP_TRY(hipMemcpyAsync(hh.data(), d_h, hh.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
      FH_HIP_TRY(hipMemcpyAsync(vn.data(), d_vnrm, V * sizeof(double), hipMemcpyDeviceToHost, stream));
      FH_HIP_TRY(hipStreamSynchronize(stream));
      for (int v = 0; v < V; ++v) {
        if (done[v]) continue;
        cd *Hv = H.data() + static_cast<size_t>(v) * (R1 + 1) * R1;  // column-major, ld R1 + 1
        cd *gv = g.data() + static_cast<size_t>(v) * (R1 + 1);
        double *cv = cs.data() + static_cast<size_t>(v) * R1;
        cd *sv = sn.data() + static_cast<size_t>(v) * R1;
        for (int i = 0; i <= j; ++i) {
          const size_t a = 2 * (static_cast<size_t>(v) * hld + i);
          const size_t b = a + 2 * static_cast<size_t>(V) * hld;
          Hv[static_cast<size_t>(j) * (R1 + 1) + i] = cd(hh[a] + hh[b], hh[a + 1] + hh[b + 1]);
        }
        const double hn = std::sqrt(vn[v]);
        Hv[static_cast<size_t>(j) * (R1 + 1) + j + 1] = cd(hn, 0.0);
        for (int i = 0; i < j; ++i) {  // earlier rotations
          cd &x0 = Hv[static_cast<size_t>(j) * (R1 + 1) + i];
          cd &x1 = Hv[static_cast<size_t>(j) * (R1 + 1) + i + 1];
          const cd t = cv[i] * x0 + sv[i] * x1;
          x1 = -std::conj(sv[i]) * x0 + cv[i] * x1;
          x0 = t;
        }
        cd &a0 = Hv[static_cast<size_t>(j) * (R1 + 1) + j];
        cd &b0 = Hv[static_cast<size_t>(j) * (R1 + 1) + j + 1];
        const double na = std::abs(a0), r = std::hypot(na, std::abs(b0));
        if (r == 0.0 || !std::isfinite(r)) {
          done[v] = 1;
          continue;
        }
        const cd ph = na == 0.0 ? cd(1.0, 0.0) : a0 / na;
        cv[j] = na / r;
        sv[j] = ph * std::conj(b0) / r;
        a0 = ph * r;
        b0 = 0.0;
        gv[j + 1] = -std::conj(sv[j]) * gv[j];
        gv[j] = cv[j] * gv[j];
        used[v] = j + 1;
        if (std::abs(gv[j + 1]) <= kGmresTol * beta[v] || hn == 0.0) done[v] = 1;
      }
    }
    steps_out = j;
    // y_v = H_v^-1 g_v (upper triangular), then DX_v = sum_i y_i b_i
    std::vector<double> y(2 * static_cast<size_t>(V) * hld, 0.0);
    for (int v = 0; v < V; ++v) {
      const int k = used[v];
      const cd *Hv = H.data() + static_cast<size_t>(v) * (R1 + 1) * R1;
      const cd *gv = g.data() + static_cast<size_t>(v) * (R1 + 1);
      std::vector<cd> yy(k);
      for (int i = k - 1; i >= 0; --i) {
        cd acc = gv[i];
        for (int c = i + 1; c < k; ++c) acc -= Hv[static_cast<size_t>(c) * (R1 + 1) + i] * yy[c];
        yy[i] = acc / Hv[static_cast<size_t>(i) * (R1 + 1) + i];
      }
      for (int i = 0; i < k; ++i) {
        y[2 * (static_cast<size_t>(v) * hld + i)] = yy[i].real();
        y[2 * (static_cast<size_t>(v) * hld + i) + 1] = yy[i].imag();
      }
    }
    FH_HIP_TRY(hipMemcpyAsync(d_y, y.data(), y.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    basis_combine_kernel<<<res_blocks, 256, 0, stream>>>(m, V, std::max(j, 1), basis, static_cast<int64_t>(n),
                                                         d_y, hld, DX, +1);
    FH_HIP_TRY(hipGetLastError());
    FH_HIP_TRY(hipStreamSynchronize(stream));  // y is read before it goes out of scope
    return FH_OK;
  };

  // Acceptance.  With e = x* - x the error of x and r = b - Zee x = Zee e,
  // the corrected Z = Zpp - b^T x - x^T r misses Z* by exactly -r^T e, so
  // |dZ|_F <= |E|_F |R|_F.  While the iteration contracts by
  // rho = |dx_k| / |dx_(k-1)| <= 1/2, |e_k| <= rho / (1 - rho) |dx_k|.
  // Converged when that bound is below kZTol relative to |Z|_F (100x under
  // the 1e-10 target), or when the first-order term |x^T r| / |Z| is below
  // kTol.
  const double kTol = 1e-13, kZTol = 1e-12;
  out.singular.assign(K, 0);
  out.refine_steps.assign(K, 0);
  out.refine_change.assign(K, INFINITY);
  out.converged.assign(K, 0);
  out.stage.assign(K, -1);
  std::vector<double> first_dx(K, 0.0);

  // Refinement steps from the current X with the current factor, classical
  // (dx = M^-1 r) or GMRES-IR; marks the frequencies that converge.
  // prev / prev_dx: the first-order term and |dx| one step earlier, carried
  // into a stage that continues after `first_it - 1` steps already taken.
  std::vector<double> prev, prev_dx;
  auto refine = [&](bool gmres, int stage, int max_steps, int first_it) -> fh_status {
    if (first_it == 1) {
      prev.assign(K, INFINITY);
      prev_dx.assign(K, 0.0);
    }
    fh_status s2;
    // frequencies that leave this stage early (their contraction is too
    // slow): each frequency's steps depend on its own history only
    std::vector<char> left(K, 0);
    for (int it = first_it; it <= max_steps; ++it) {
      int running = 0;
      for (int f = 0; f < K; ++f) {
        h_active[f] = out.converged[f] || left[f] ? 0 : 1;
        running += h_active[f];
      }
      if (running == 0) break;
      FH_HIP_TRY(hipMemcpyAsync(d_active, h_active.data(), K * sizeof(int), hipMemcpyHostToDevice, stream));
      if (gmres) {
        int inner = 0;
        if ((s2 = gmres_correction(inner))) return s2;
      } else {
        axpy_copy_kernel<<<vec_blocks, 256, 0, stream>>>(static_cast<int64_t>(n), nullptr, nullptr, DX, R);
        if ((s2 = factor_solve<float>(L, F, V, DX, P, stream))) return s2;
      }
      active_axpy_kernel<<<vec_blocks, 256, 0, stream>>>(m, p, static_cast<int64_t>(n), X, DX, d_active);
      if ((s2 = residual())) return s2;
      functional_kernel<<<K * p * p, 256, 0, stream>>>(m, p, d_omegas, in.d_rep, in.d_rpp,
                                                        in.d_lmesh, lds, X, R, d_z, metric, d_active);
      norms_kernel<<<V, 256, 0, stream>>>(m, DX, R, d_nrm);
      FH_HIP_TRY(hipGetLastError());
      FH_HIP_TRY(hipMemcpyAsync(h_metric.data(), metric, h_metric.size() * sizeof(double),
                                hipMemcpyDeviceToHost, stream));
      FH_HIP_TRY(hipMemcpyAsync(h_nrm.data(), d_nrm, h_nrm.size() * sizeof(double),
                                hipMemcpyDeviceToHost, stream));
      FH_HIP_TRY(hipMemcpyAsync(h_z.data(), d_z, h_z.size() * sizeof(double),
                                hipMemcpyDeviceToHost, stream));
      FH_HIP_TRY(hipStreamSynchronize(stream));
      bool all = true;
      for (int f = 0; f < K; ++f) {
        if (!h_active[f]) continue;
        double worst = 0.0;
        for (int q = 0; q < p * p; ++q)
          worst = std::max(worst, h_metric[static_cast<size_t>(f) * p * p + q]);
        if (!(worst == worst)) worst = INFINITY;
        out.refine_steps[f] += 1;
        out.refine_change[f] = worst;
        double sdx = 0.0, sr = 0.0, sz = 0.0;
        for (int q = 0; q < p; ++q) {
          sdx += h_nrm[2 * (static_cast<size_t>(f) * p + q)];
          sr += h_nrm[2 * (static_cast<size_t>(f) * p + q) + 1];
        }
        for (int q = 0; q < 2 * p * p; ++q) {
          const double zq = h_z[2 * static_cast<size_t>(f) * p * p + q];
          sz += zq * zq;
        }
        const double ndx = std::sqrt(sdx), nr = std::sqrt(sr), nz = std::sqrt(sz);
        double bound = INFINITY, x_bound = INFINITY;
        bool stalled = false;
        // X starts at 0: the first correction of the first stage is X
        if (first_dx[f] == 0.0 && std::isfinite(ndx)) first_dx[f] = ndx;
        if (it >= 2 && prev_dx[f] > 0.0) {
          const double rho = ndx / prev_dx[f];
          if (rho <= 0.5) x_bound = rho / (1.0 - rho) * ndx;
          if (rho <= 0.5 && nz > 0.0) bound = x_bound * nr / nz;
          stalled = it >= 3 && rho > 0.5;
        }
        prev_dx[f] = ndx;
        // Converged when the bound on the error of Z is small, when the
        // first-order term is negligible, or when that term has stopped
        // shrinking at a level the corrected Z absorbs.
        bool done = std::isfinite(nz) && h_sing[f] == 0 &&
                    (bound <= kZTol || worst <= kTol || (worst <= 1e-9 && !(worst < 0.5 * prev[f])));
        // x itself to working accuracy when asked for (it enters gradients
        // to first order): the correction's bound under 1e-14 |x|, or the
        // contraction stalled at the rounding level, or x exact.
        if (done && in.tight_x) done = ndx == 0.0 || x_bound <= 1e-14 * first_dx[f] || stalled;
        if (done) {
          out.converged[f] = 1;
          out.stage[f] = stage;
        } else {
          all = false;
        }
        prev[f] = worst;
      }
      if (all) break;
      // A contraction slower than 1/4 per step will not reach the tolerance
      // within the budget: such a frequency leaves early for the next stage.
      if (it >= 3)
        for (int f = 0; f < K; ++f)
          if (h_active[f] && !out.converged[f] && !(out.refine_change[f] < 1e-6)) left[f] = 1;
    }
    return FH_OK;
  };

  // The stages, in order: split-bf16 factor with classical refinement, the
  // same factor under GMRES-IR, the FP32 factor with classical refinement,
  // the FP32 factor under GMRES-IR (in.refine_stages selects which run).
  // The split copies of one frequency must fit the 2 GiB range of a buffer
  // descriptor (far above the dense path's 10 000 filaments).
  const bool split_ok = split_blocks(L) * 4096 < (int64_t{1} << 31);
  int factored = -1;  // 0 split, 1 FP32
  bool started = false;
  int stage0_from = 1;  // first classical step of stage 0 still to run on the host
  const unsigned kb = static_cast<unsigned>((K + 255) / 256);
  // The split factorization and its first kGraphSteps classical steps with
  // the acceptance test on the device: enqueued as they are, or captured
  // into the chunk's graph and replayed.
  auto stage0_steps = [&](Profiler *pr) -> fh_status {
    fill_int_kernel<<<kb, 256, 0, stream>>>(K, d_sing, 0);
    fh_status s2;
    if ((s2 = factorize<float>(in, L, d_omegas, K, F, d_sing, stream, pr, &sb))) return s2;
    fill_kernel<<<vec_blocks, 256, 0, stream>>>(static_cast<int64_t>(n), X, 0.0);
    refine_state_init_kernel<<<kb, 256, 0, stream>>>(K, d_state);
    // the residual of X = 0 is b itself
    border_rhs_kernel<<<res_blocks, 256, 0, stream>>>(m, p, V, d_omegas, in.d_rep, in.d_lmesh, lds, R);
    FH_HIP_TRY(hipGetLastError());
    for (int it = 1; it <= kGraphSteps; ++it) {
      axpy_copy_kernel<<<vec_blocks, 256, 0, stream>>>(static_cast<int64_t>(n), nullptr, nullptr, DX, R);
      if ((s2 = factor_solve<float>(L, F, V, DX, P, stream))) return s2;
      axpy_copy_kernel<<<vec_blocks, 256, 0, stream>>>(static_cast<int64_t>(n), X, DX, nullptr, nullptr);
      if ((s2 = residual())) return s2;
      functional_kernel<<<K * p * p, 256, 0, stream>>>(m, p, d_omegas, in.d_rep, in.d_rpp,
                                                        in.d_lmesh, lds, X, R, d_z, metric, nullptr);
      norms_kernel<<<V, 256, 0, stream>>>(m, DX, R, d_nrm);
      refine_accept_kernel<<<kb, 256, 0, stream>>>(K, p, it, metric, d_nrm, d_z, d_sing, kZTol, kTol,
                                                   d_state);
      FH_HIP_TRY(hipGetLastError());
    }
    return FH_OK;
  };
  if (in.graph && !in.tight_x && (in.refine_stages & 1u) && split_ok) {
    RefineGraph &gr = *in.graph;
    const void *mats[4] = {in.d_lmesh, in.d_ree_val, in.d_rep, in.d_r_vals};
    const bool same = gr.base == ws.base && gr.omegas == d_omegas && gr.z == d_z && gr.K == K &&
                      gr.nb == L.nb && gr.profiled == (prof != nullptr) &&
                      std::equal(mats, mats + 4, gr.mats);
    if (!same) {
      gr.reset();
      std::copy(mats, mats + 4, gr.mats);
      gr.base = ws.base;
      gr.omegas = d_omegas;
      gr.z = d_z;
      gr.K = K;
      gr.nb = L.nb;
      gr.profiled = prof != nullptr;
    }
    ++gr.seen;
    if (gr.exec || gr.seen >= 2) {
      if (!gr.exec) {
        FH_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        gr.prof.captured = true;
        st = stage0_steps(gr.profiled ? &gr.prof : nullptr);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(stream, &g);
        if (st) {
          if (g) (void)hipGraphDestroy(g);
          gr.reset();
          return st;
        }
        if (e != hipSuccess || !g) {
          gr.reset();
          return set_error(FH_E_HIP, std::string("graph capture of the refined solve failed: ") +
                                         hipGetErrorString(e));
        }
        gr.graph = g;
        FH_HIP_TRY(hipGraphInstantiate(&gr.exec, gr.graph, nullptr, nullptr, 0));
      }
      FH_HIP_TRY(hipGraphLaunch(gr.exec, stream));
    } else {
      if ((st = stage0_steps(prof))) return st;
    }
  } else if ((in.refine_stages & 1u) && split_ok && !in.tight_x) {
    if ((st = stage0_steps(prof))) return st;
  }
  if ((in.refine_stages & 1u) && split_ok && !in.tight_x) {
    // the device state of the first steps, into the host's bookkeeping
    std::vector<double> hs(6 * static_cast<size_t>(K));
    FH_HIP_TRY(hipMemcpyAsync(hs.data(), d_state, hs.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    FH_HIP_TRY(hipMemcpyAsync(h_sing.data(), d_sing, K * sizeof(int), hipMemcpyDeviceToHost, stream));
    FH_HIP_TRY(hipStreamSynchronize(stream));
    if (in.graph && in.graph->exec && prof) prof->absorb(in.graph->prof);
    prev.assign(K, INFINITY);
    prev_dx.assign(K, 0.0);
    for (int f = 0; f < K; ++f) {
      out.converged[f] = hs[f] != 0.0 ? 1 : 0;
      out.stage[f] = out.converged[f] ? 0 : -1;
      out.refine_steps[f] = static_cast<int>(hs[K + f]);
      out.refine_change[f] = hs[2 * K + f];
      prev[f] = hs[3 * K + f];
      prev_dx[f] = hs[4 * K + f];
      first_dx[f] = hs[5 * K + f];
    }
    factored = 0;
    started = true;
    out.factor_arith = 2;
    stage0_from = kGraphSteps + 1;
  }
  for (int stage = 0; stage < 4; ++stage) {
    if (!(in.refine_stages & (1u << stage))) continue;
    const int want = stage < 2 ? 0 : 1;
    if (want == 0 && !split_ok) continue;
    bool all = true;
    for (int f = 0; f < K; ++f) all = all && out.converged[f];
    if (all) break;
    if (factored != want) {
      fill_int_kernel<<<kb, 256, 0, stream>>>(K, d_sing, 0);
      if ((st = factorize<float>(in, L, d_omegas, K, F, d_sing, stream, prof, want == 0 ? &sb : nullptr)))
        return st;
      FH_HIP_TRY(hipMemcpyAsync(h_sing.data(), d_sing, K * sizeof(int), hipMemcpyDeviceToHost, stream));
      factored = want;
      out.factor_arith = want == 0 ? 2 : 1;
      if (!started) {
        fill_kernel<<<vec_blocks, 256, 0, stream>>>(static_cast<int64_t>(n), X, 0.0);
        started = true;
      } else {
        // restart the frequencies not yet converged from x = 0: an earlier
        // factor may have left them non-finite
        for (int f = 0; f < K; ++f) h_reset[f] = out.converged[f] ? 0 : 1;
        FH_HIP_TRY(hipMemcpyAsync(d_reset, h_reset.data(), K * sizeof(int), hipMemcpyHostToDevice, stream));
        reset_vectors_kernel<<<vec_blocks, 256, 0, stream>>>(m, p, V, d_reset, X);
        FH_HIP_TRY(hipGetLastError());
        for (int f = 0; f < K; ++f)
          if (!out.converged[f]) first_dx[f] = 0.0;
      }
      if ((st = residual())) return st;
      FH_HIP_TRY(hipStreamSynchronize(stream));  // h_sing, h_reset
    }
    const bool gmres = stage == 1 || stage == 3;
    if ((st = refine(gmres, stage, gmres ? kMaxGmresRefine : kMaxRefine, stage == 0 ? stage0_from : 1)))
      return st;
  }
  // A zero pivot of an FP32 factor is not a singular system: it is reported
  // as not converged, for the caller to solve in FP64.
  if (in.d_x_out)
    FH_HIP_TRY(hipMemcpyAsync(in.d_x_out, X, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
  FH_HIP_TRY(hipStreamSynchronize(stream));
  return FH_OK;
}

}  // namespace fh
