// Pruned FFT convolution on the compact pFFT grid (fasterhenry/src/pfft/
// fft.rs:91-217), with hand-written LDS line FFTs.
//
// A "line pass" transforms lines of a 3-D array (a, l, b) along l: element
// (a, l, b) sits at (a L + l) B + b, so one kernel serves every axis.  A
// workgroup takes X consecutive lines (consecutive b, so the loads and
// stores of one l are X contiguous complex values; for B = 1 the lines
// themselves are contiguous), zero-pads them from L_in to the padded length
// M in LDS, runs in-place mixed-radix (4, 2, 3, 5) Cooley-Tukey passes
// (forward: digit-reversed load, decimation in time; inverse: decimation in
// frequency, digit-reversed store), and stores the first L_out points.
//
// Convolution (sources in [c][N0][N1][N2]):
//   axis 2 forward  [c][N0][N1][N2]  -> [c][N0][N1][M2]   (lines contiguous)
//   axis 1 forward  [c][N0][N1][M2]  -> [c][N0][M1][M2]
//   axis 0 forward, times the spectrum, axis 0 inverse, in place on the
//                   x < N0 rows of [c][N0][M1][M2]
//   axis 1 inverse  [c][N0][M1][M2]  -> [c][N0][N1][M2]
//   axis 2 inverse  [c][N0][N1][M2]  -> [c][N0][N1][N2]
// Only N0 of the M0 axis-0 planes and N0 N1 of the axis-2 columns are ever
// touched: the zero padding is never stored or read.
//
// The spectrum of the wrapped kernel 1/(h |d|) is real and even in every
// axis; it is computed once in FP64 with the same line passes on the full
// padded grid and kept for k_a <= M_a / 2 (1/8 of the padded grid), scaled
// by 1/(M0 M1 M2) since the inverse passes are unnormalized.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "fh_fft.hpp"

namespace fh {
namespace {

constexpr int kThreads = 256;       // element-wise kernels
constexpr int kLineThreads = 1024;  // line passes: 16 waves share one LDS copy of the lines
constexpr int kMaxPasses = 16;
constexpr int kMaxLines = 256;
constexpr int kLdsBudget = 150 * 1024;  // dynamic LDS of a line pass (lines + twiddles)
constexpr int kMaxLine = 8191;          // longest padded axis of the line passes
constexpr int kDirectMax = 32;          // axis-2 lengths transformed by a direct DFT

// n / d for n, d < 2^16 as umulhi(n, ceil(2^32 / d)) (exact in that range:
// the error term n (m d - 2^32) / 2^32 stays below 1 / d).
__host__ __device__ inline uint32_t magic_of(uint32_t d) {
  return d <= 1 ? 0u : static_cast<uint32_t>((0xFFFFFFFFull / d) + 1);
}
__device__ inline int fast_div(int n, int d, uint32_t m) {
  return d == 1 ? n : static_cast<int>(__umulhi(static_cast<uint32_t>(n), m));
}

struct LinePlan {
  int M;
  int npass;
  int radix[kMaxPasses];
  uint32_t span_magic[kMaxPasses];  // magic of Ns for pass p
  uint32_t nb_magic[kMaxPasses];    // magic of M / radix[p]
  uint32_t m_magic;
};

bool make_plan(int64_t M, LinePlan &p) {
  p.M = static_cast<int>(M);
  p.npass = 0;
  int64_t r = M;
  auto take = [&](int f) {
    while (r % f == 0 && p.npass < kMaxPasses) {
      p.radix[p.npass++] = f;
      r /= f;
    }
  };
  take(4);
  take(2);
  take(3);
  take(5);
  int64_t span = 1;
  for (int q = 0; q < p.npass; ++q) {
    p.span_magic[q] = magic_of(static_cast<uint32_t>(span));
    p.nb_magic[q] = magic_of(static_cast<uint32_t>(M / p.radix[q]));
    span *= p.radix[q];
  }
  p.m_magic = magic_of(static_cast<uint32_t>(M));
  return r == 1;
}

template <typename T2>
struct Real;
template <>
struct Real<float2> {
  using type = float;
};
template <>
struct Real<double2> {
  using type = double;
};

__device__ inline float2 cx(float a, float b) { return make_float2(a, b); }
__device__ inline double2 cx(double a, double b) { return make_double2(a, b); }

template <typename T2>
__device__ inline T2 cadd(T2 a, T2 b) {
  return cx(a.x + b.x, a.y + b.y);
}
template <typename T2>
__device__ inline T2 csub(T2 a, T2 b) {
  return cx(a.x - b.x, a.y - b.y);
}
template <typename T2>
__device__ inline T2 cmul(T2 a, T2 b) {
  return cx(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// i s z for s = +-1
template <int S, typename T2>
__device__ inline T2 irot(T2 z) {
  return S > 0 ? cx(-z.y, z.x) : cx(z.y, -z.x);
}
template <typename T2, typename R>
__device__ inline T2 cscale(T2 a, R s) {
  return cx(a.x * s, a.y * s);
}

// In-register DFT of R points with kernel exp(S 2 pi i k n / R).
template <int R, int S, typename T2>
__device__ inline void dft(T2 *v) {
  using Re = typename Real<T2>::type;
  if constexpr (R == 2) {
    const T2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 3) {
    const Re h = static_cast<Re>(-0.5), s = static_cast<Re>(0.86602540378443864676);
    const T2 t = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    const T2 m = cadd(v[0], cscale(t, h));
    const T2 r = irot<S>(cscale(d, s));
    v[0] = cadd(v[0], t);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  } else if constexpr (R == 4) {
    const T2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    const T2 c = cadd(v[1], v[3]), d = irot<S>(csub(v[1], v[3]));
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, d);
    v[3] = csub(b, d);
  } else {
    const Re c1 = static_cast<Re>(0.30901699437494742410), c2 = static_cast<Re>(-0.80901699437494742410);
    const Re s1 = static_cast<Re>(0.95105651629515357212), s2 = static_cast<Re>(0.58778525229247312917);
    const T2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    const T2 d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
    const T2 m1 = cadd(v[0], cadd(cscale(t1, c1), cscale(t2, c2)));
    const T2 m2 = cadd(v[0], cadd(cscale(t1, c2), cscale(t2, c1)));
    const T2 r1 = irot<S>(cadd(cscale(d1, s1), cscale(d2, s2)));
    const T2 r2 = irot<S>(csub(cscale(d1, s2), cscale(d2, s1)));
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = cadd(m1, r1);
    v[4] = csub(m1, r1);
    v[2] = cadd(m2, r2);
    v[3] = csub(m2, r2);
  }
}

// In-place Cooley-Tukey passes over `lines` lines of length M in LDS (line
// stride P).  Pass s of radix R = r_s has span Ns = r_1 ... r_(s-1); its
// butterfly (group, k), k < Ns, owns the slots group Ns R + k + t Ns,
// t < R, and writes its outputs back to them, so one LDS copy and one
// barrier per pass suffice.  `tw` holds exp(-2 pi i q / M); the factor of
// slot t is w^(k t) with w = exp(S 2 pi i / (Ns R)).
//  DIT (twiddle, then DFT_R), passes 1 .. p: digit-reversed input ->
//  natural-order output.  DIF (DFT_R, then twiddle), passes p .. 1: the
//  transpose of the DIT product, natural input -> digit-reversed output.
template <typename T2, int R, int S, bool DIT>
__device__ void ct_pass(T2 *lds, int P, int M, int lines, int Ns, uint32_t ns_magic, uint32_t nb_magic,
                        const T2 *tw) {
  const int nb = M / R, total = nb * lines;
  const int step = M / (Ns * R);
  for (int q = threadIdx.x; q < total; q += kLineThreads) {
    const int line = fast_div(q, nb, nb_magic), b = q - line * nb;
    const int grp = fast_div(b, Ns, ns_magic), k = b - grp * Ns;
    T2 *base = lds + line * P + grp * Ns * R + k;
    T2 v[R];
#pragma unroll
    for (int t = 0; t < R; ++t) v[t] = base[t * Ns];
    if (DIT) {
#pragma unroll
      for (int t = 1; t < R; ++t) {
        T2 w = tw[k * t * step];
        if (S > 0) w.y = -w.y;
        v[t] = cmul(v[t], w);
      }
    }
    dft<R, S>(v);
    if (!DIT) {
#pragma unroll
      for (int t = 1; t < R; ++t) {
        T2 w = tw[k * t * step];
        if (S > 0) w.y = -w.y;
        v[t] = cmul(v[t], w);
      }
    }
#pragma unroll
    for (int t = 0; t < R; ++t) base[t * Ns] = v[t];
  }
  __syncthreads();
}

template <typename T2, int S, bool DIT>
__device__ inline void ct_pass_r(const LinePlan &plan, int p, T2 *lds, int P, int lines, int Ns, const T2 *tw) {
  const int r = plan.radix[p], M = plan.M;
  const uint32_t sm = plan.span_magic[p], nm = plan.nb_magic[p];
  if (r == 4) ct_pass<T2, 4, S, DIT>(lds, P, M, lines, Ns, sm, nm, tw);
  else if (r == 2) ct_pass<T2, 2, S, DIT>(lds, P, M, lines, Ns, sm, nm, tw);
  else if (r == 3) ct_pass<T2, 3, S, DIT>(lds, P, M, lines, Ns, sm, nm, tw);
  else ct_pass<T2, 5, S, DIT>(lds, P, M, lines, Ns, sm, nm, tw);
}

// forward DFT (S = -1) by DIT: digit-reversed in, natural out
template <typename T2>
__device__ void forward_dit(T2 *lds, int P, int lines, const LinePlan &plan, const T2 *tw) {
  int Ns = 1;
  for (int p = 0; p < plan.npass; ++p) {
    ct_pass_r<T2, -1, true>(plan, p, lds, P, lines, Ns, tw);
    Ns *= plan.radix[p];
  }
}

// inverse DFT (S = +1, unnormalized) by DIF: natural in, digit-reversed out
template <typename T2>
__device__ void inverse_dif(T2 *lds, int P, int lines, const LinePlan &plan, const T2 *tw) {
  int Ns = plan.M;
  for (int p = plan.npass - 1; p >= 0; --p) {
    Ns /= plan.radix[p];
    ct_pass_r<T2, 1, false>(plan, p, lds, P, lines, Ns, tw);
  }
}

// Spectrum layout for the fused axis-0 pass: line b = k1 M2 + k2, value at
// k0 is spec[(k2' H1 + k1') H0 + k0'] with k' = min(k, M - k), H = M/2 + 1.
struct SpecDesc {
  int64_t M1, M2, H0, H1;
};

// mode 0: forward; 1: inverse; 2: forward, times the spectrum, inverse.
// Workgroup w of the launch takes line group (w mod 8) G8 + w / 8 (G8 =
// groups per XCD), so neighbouring line groups -- whose strided loads share
// cache lines -- run on one XCD and meet in its L2.  TW_LDS: the twiddle
// table is copied behind the lines in LDS.
template <typename T2, bool TW_LDS>
__global__ void __launch_bounds__(kLineThreads) line_pass_kernel(
    const T2 *in, T2 *out, int64_t lines_total, int64_t groups, int64_t B, int L_in, int L_out, LinePlan plan,
    const T2 *__restrict__ tw_global, const int *__restrict__ rev, int X, int P, int mode,
    const typename Real<T2>::type *__restrict__ spec, SpecDesc sd) {
  extern __shared__ __align__(16) unsigned char smem[];
  T2 *lds = reinterpret_cast<T2 *>(smem);
  __shared__ int64_t base_in[kMaxLines], base_out[kMaxLines], spec_base[kMaxLines];
  const int64_t per_xcd = (groups + 7) / 8;
  const int64_t group = static_cast<int64_t>(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (group >= groups) return;
  const int64_t first = group * X;
  const int lines = static_cast<int>(min(static_cast<int64_t>(X), lines_total - first));
  const int M = plan.M;
  const T2 *tw = tw_global;
  if (TW_LDS) {
    T2 *t = lds + X * P;
    for (int q = threadIdx.x; q < M; q += kLineThreads) t[q] = tw_global[q];
    tw = t;
  }
  for (int xi = threadIdx.x; xi < lines; xi += kLineThreads) {
    const int64_t lam = first + xi, a = lam / B, b = lam - a * B;
    base_in[xi] = a * L_in * B + b;
    base_out[xi] = a * L_out * B + b;
    if (mode == 2) {
      const int64_t k1 = b / sd.M2, k2 = b - k1 * sd.M2;
      const int64_t r1 = min(k1, sd.M1 - k1), r2 = min(k2, sd.M2 - k2);
      spec_base[xi] = (r2 * sd.H1 + r1) * sd.H0;
    }
  }
  __syncthreads();
  const uint32_t lines_magic = magic_of(static_cast<uint32_t>(lines));
  const int elems = lines * M;
  for (int idx = threadIdx.x; idx < elems; idx += kLineThreads) {
    int xi, l;
    if (B == 1) {
      xi = fast_div(idx, M, plan.m_magic);
      l = idx - xi * M;
    } else {
      l = fast_div(idx, lines, lines_magic);
      xi = idx - l * lines;
    }
    T2 v = cx(static_cast<typename Real<T2>::type>(0), static_cast<typename Real<T2>::type>(0));
    if (l < L_in) v = in[base_in[xi] + static_cast<int64_t>(l) * B];
    lds[xi * P + (mode == 1 ? l : rev[l])] = v;
  }
  __syncthreads();
  if (mode != 1) forward_dit<T2>(lds, P, lines, plan, tw);
  if (mode == 2) {
    for (int idx = threadIdx.x; idx < elems; idx += kLineThreads) {
      const int xi = fast_div(idx, M, plan.m_magic), l = idx - xi * M;
      const int r0 = min(l, M - l);
      lds[xi * P + l] = cscale(lds[xi * P + l], spec[spec_base[xi] + r0]);
    }
    __syncthreads();
  }
  if (mode != 0) inverse_dif<T2>(lds, P, lines, plan, tw);
  const int oelems = lines * L_out;
  const uint32_t lout_magic = magic_of(static_cast<uint32_t>(L_out));
  for (int idx = threadIdx.x; idx < oelems; idx += kLineThreads) {
    int xi, l;
    if (B == 1) {
      xi = fast_div(idx, L_out, lout_magic);
      l = idx - xi * L_out;
    } else {
      l = fast_div(idx, lines, lines_magic);
      xi = idx - l * lines;
    }
    out[base_out[xi] + static_cast<int64_t>(l) * B] = lds[xi * P + (mode == 0 ? l : rev[l])];
  }
}

// Axis-2 transforms of short padded length (M2 <= 32) as direct DFTs, one
// thread per column: forward N2 -> M2 points, inverse M2 -> N2 points.
// The workgroup's 256 output columns are one contiguous stretch of `out`:
// the results go through LDS (column stride L_out | 1, odd, so the 8- and
// 16-byte writes of neighbouring threads fall in different banks) and are
// stored in order, whole cache lines, instead of one element per thread at
// a stride of L_out elements.  In place (in == out) needs L_in == L_out:
// every workgroup then reads and writes only its own stretch.
template <typename T2>
__global__ void __launch_bounds__(kThreads) column_dft_kernel(const T2 *in, T2 *out,
                                                                int64_t cols, int L_in, int L_out, int M,
                                                                const T2 *__restrict__ tw_global, int inverse) {
  using Re = typename Real<T2>::type;
  __shared__ T2 tw[kDirectMax];
  extern __shared__ __align__(16) unsigned char smem[];
  T2 *st = reinterpret_cast<T2 *>(smem);  // [kThreads][L_out | 1]
  const int P = L_out | 1;
  if (threadIdx.x < M) tw[threadIdx.x] = tw_global[threadIdx.x];
  __syncthreads();
  const int64_t c0 = blockIdx.x * static_cast<int64_t>(kThreads);
  const int64_t c = c0 + threadIdx.x;
  if (c < cols) {
    T2 v[kDirectMax];
#pragma unroll
    for (int q = 0; q < kDirectMax; ++q)
      v[q] = q < L_in ? in[c * L_in + q] : cx(static_cast<Re>(0), static_cast<Re>(0));
    for (int k = 0; k < L_out; ++k) {
      T2 acc = cx(static_cast<Re>(0), static_cast<Re>(0));
      int e = 0;  // (k q) mod M, advanced by k per step
#pragma unroll
      for (int q = 0; q < kDirectMax; ++q) {
        if (q < L_in) {
          T2 w = tw[e];
          if (inverse) w.y = -w.y;
          acc = cadd(acc, cmul(v[q], w));
          e += k;
          if (e >= M) e -= M;
        }
      }
      st[threadIdx.x * P + k] = acc;
    }
  }
  __syncthreads();
  const int ncol = static_cast<int>(min(static_cast<int64_t>(kThreads), cols - c0));
  const int total = ncol * L_out;
  T2 *dst = out + c0 * L_out;
  for (int e = threadIdx.x; e < total; e += kThreads) {
    const int cc = e / L_out, k = e - cc * L_out;
    dst[e] = st[cc * P + k];
  }
}

// The kernel 1/(h |d|) on the padded grid at wrapped offsets (fft.rs:101-128).
__global__ void wrapped_kernel_kernel(double2 *__restrict__ data, int64_t n0, int64_t n1, int64_t n2,
                                      int64_t m0, int64_t m1, int64_t m2, double h) {
  const int64_t total = m0 * m1 * m2;
  for (int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; q < total;
       q += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t k = q % m2, j = (q / m2) % m1, i = q / (m2 * m1);
    bool ok = true;
    auto off = [&](int64_t x, int64_t n, int64_t m) -> int64_t {
      if (x < n) return x;
      if (x + n > m) return x - m;
      ok = false;
      return 0;
    };
    const int64_t di = off(i, n0, m0), dj = off(j, n1, m1), dk = off(k, n2, m2);
    data[q] = make_double2(ok ? pfft_grid_kernel(di, dj, dk, h) : 0.0, 0.0);
  }
}

template <typename Re>
__global__ void half_spectrum_kernel(const double2 *__restrict__ data, Re *__restrict__ spec, int64_t m1,
                                     int64_t m2, int64_t h0, int64_t h1, int64_t h2, double scale) {
  const int64_t total = h0 * h1 * h2;
  for (int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; q < total;
       q += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t k0 = q % h0, k1 = (q / h0) % h1, k2 = q / (h0 * h1);
    spec[q] = static_cast<Re>(data[(k0 * m1 + k1) * m2 + k2].x * scale);
  }
}

unsigned blocks_for(int64_t total) {
  return static_cast<unsigned>(std::min<int64_t>((total + kThreads - 1) / kThreads, 1 << 20));
}

}  // namespace

struct ConvolverImpl {
  bool single = false;
  int64_t N[3] = {0, 0, 0}, M[3] = {0, 0, 0}, H[3] = {0, 0, 0};
  LinePlan plan[3];
  int64_t tw_off[3] = {0, 0, 0};
  int max_channels = 0;
  DeviceArray<char> tw32, tw64, grid, bufA, bufB, spec;
  DeviceArray<int> rev;
  // lines per workgroup of the convolution's line passes, by (axis, mode);
  // 0 = the largest that fits the LDS budget
  int xsel[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};

  template <typename T2>
  fh_status pass(const T2 *in, T2 *out, int64_t lines, int64_t B, int L_in, int L_out, int axis, int mode,
                 const T2 *tw, hipStream_t stream, int x_override = -1) {
    if (lines <= 0) return FH_OK;
    const int Mx = plan[axis].M;
    if (axis == 2 && mode != 2 && Mx <= kDirectMax && B == 1) {
      // in place only column for column (each workgroup's stretch is its own)
      if (in == out && L_in != L_out)
        return set_error(FH_E_PFFT, "axis-2 direct DFT in place with unequal lengths", 0, 0, 0.0, 2);
      const int64_t blocks = (lines + kThreads - 1) / kThreads;
      if (blocks > 0x7fffffff) return set_error(FH_E_PFFT, "pFFT grid too large for the axis-2 DFT", 0, 0, 0.0, 2);
      const size_t lds = static_cast<size_t>(kThreads) * (L_out | 1) * sizeof(T2);
      column_dft_kernel<T2><<<static_cast<unsigned>(blocks), kThreads, lds, stream>>>(
          in, out, lines, L_in, L_out, Mx, tw + tw_off[axis], mode == 1 ? 1 : 0);
      FH_HIP_TRY(hipGetLastError());
      return FH_OK;
    }
    const int P = Mx + 1;
    const size_t esz = sizeof(T2);
    const bool tw_lds = static_cast<size_t>(P + Mx) * esz <= static_cast<size_t>(kLdsBudget);
    const size_t room = kLdsBudget - (tw_lds ? Mx * esz : 0);
    int X = static_cast<int>(std::max<size_t>(1, std::min<size_t>(kMaxLines, room / (P * esz))));
    if (B > 1) {  // whole powers of two of lines: aligned segments of the strided loads
      int p2 = 1;
      while (p2 * 2 <= X) p2 *= 2;
      X = p2;
    }
    const int chosen = x_override >= 0 ? x_override : xsel[axis][mode];
    if (chosen > 0) X = std::min(X, chosen);
    X = static_cast<int>(std::min<int64_t>(X, lines));
    const size_t lds = (static_cast<size_t>(X) * P + (tw_lds ? Mx : 0)) * esz;
    const int64_t groups = (lines + X - 1) / X;
    const int64_t blocks = (groups + 7) / 8 * 8;
    if (blocks > 0x7fffffff) return set_error(FH_E_PFFT, "pFFT grid too large for the line passes", 0, 0, 0.0, 2);
    SpecDesc sd{M[1], M[2], H[0], H[1]};
    using Re = typename Real<T2>::type;
    if (tw_lds)
      line_pass_kernel<T2, true><<<static_cast<unsigned>(blocks), kLineThreads, lds, stream>>>(
          in, out, lines, groups, B, L_in, L_out, plan[axis], tw + tw_off[axis], rev.ptr + tw_off[axis], X, P,
          mode, reinterpret_cast<const Re *>(spec.ptr), sd);
    else
      line_pass_kernel<T2, false><<<static_cast<unsigned>(blocks), kLineThreads, lds, stream>>>(
          in, out, lines, groups, B, L_in, L_out, plan[axis], tw + tw_off[axis], rev.ptr + tw_off[axis], X, P,
          mode, reinterpret_cast<const Re *>(spec.ptr), sd);
    FH_HIP_TRY(hipGetLastError());
    return FH_OK;
  }

  template <typename T2>
  fh_status convolve_t(int C, const T2 *tw, hipStream_t stream) {
    T2 *g = reinterpret_cast<T2 *>(grid.ptr), *a = reinterpret_cast<T2 *>(bufA.ptr),
       *b = reinterpret_cast<T2 *>(bufB.ptr);
    const int n0 = static_cast<int>(N[0]), n1 = static_cast<int>(N[1]), n2 = static_cast<int>(N[2]);
    const int m1 = static_cast<int>(M[1]), m2 = static_cast<int>(M[2]);
    fh_status st;
    if ((st = pass<T2>(g, a, C * N[0] * N[1], 1, n2, m2, 2, 0, tw, stream)) ||
        (st = pass<T2>(a, b, C * N[0] * M[2], M[2], n1, m1, 1, 0, tw, stream)) ||
        (st = pass<T2>(b, b, C * M[1] * M[2], M[1] * M[2], n0, n0, 0, 2, tw, stream)) ||
        (st = pass<T2>(b, a, C * N[0] * M[2], M[2], m1, n1, 1, 1, tw, stream)) ||
        (st = pass<T2>(a, g, C * N[0] * N[1], 1, m2, n2, 2, 1, tw, stream)))
      return st;
    return FH_OK;
  }

  // The largest X of a pass (the LDS rule) and its power-of-two divisors.
  int max_lines(int axis, size_t esz) const {
    const int Mx = plan[axis].M, P = Mx + 1;
    const bool tw_lds = static_cast<size_t>(P + Mx) * esz <= static_cast<size_t>(kLdsBudget);
    const size_t room = kLdsBudget - (tw_lds ? Mx * esz : 0);
    int X = static_cast<int>(std::max<size_t>(1, std::min<size_t>(kMaxLines, room / (P * esz))));
    int p2 = 1;
    while (p2 * 2 <= X) p2 *= 2;
    return p2;
  }

  // Times each candidate lines-per-workgroup of the three long-line passes
  // on the operator's own buffers and keeps the fastest.  Every line is
  // transformed by the same operations whatever X is, so the choice changes
  // speed only.
  template <typename T2>
  fh_status tune(const T2 *tw, hipStream_t stream) {
    T2 *a = reinterpret_cast<T2 *>(bufA.ptr), *b = reinterpret_cast<T2 *>(bufB.ptr);
    const int C = std::min(2, max_channels);
    const int n0 = static_cast<int>(N[0]), n1 = static_cast<int>(N[1]);
    const int m1 = static_cast<int>(M[1]);
    hipEvent_t e0, e1;
    FH_HIP_TRY(hipEventCreate(&e0));
    FH_HIP_TRY(hipEventCreate(&e1));
    struct Guard {
      hipEvent_t a, b;
      ~Guard() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
      }
    } guard{e0, e1};
    struct Stage {
      int axis, mode;
    } stages[3] = {{1, 0}, {0, 2}, {1, 1}};
    for (const Stage &sg : stages) {
      const int xmax = max_lines(sg.axis, sizeof(T2));
      float best = 0.0f;
      int best_x = 0;
      for (int x = xmax; x >= 1 && x * 8 >= xmax; x /= 2) {
        auto run = [&]() -> fh_status {
          if (sg.axis == 1 && sg.mode == 0) return pass<T2>(a, b, C * N[0] * M[2], M[2], n1, m1, 1, 0, tw, stream, x);
          if (sg.axis == 0) return pass<T2>(b, b, C * M[1] * M[2], M[1] * M[2], n0, n0, 0, 2, tw, stream, x);
          return pass<T2>(b, a, C * N[0] * M[2], M[2], m1, n1, 1, 1, tw, stream, x);
        };
        fh_status st;
        if ((st = run())) return st;
        FH_HIP_TRY(hipEventRecord(e0, stream));
        for (int rep = 0; rep < 3; ++rep)
          if ((st = run())) return st;
        FH_HIP_TRY(hipEventRecord(e1, stream));
        FH_HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.0f;
        FH_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        if (best_x == 0 || ms < best) {
          best = ms;
          best_x = x;
        }
      }
      xsel[sg.axis][sg.mode] = best_x;
    }
    return FH_OK;
  }
};

PrunedConvolver::PrunedConvolver() : impl_(new ConvolverImpl) {}
PrunedConvolver::~PrunedConvolver() { delete impl_; }
void *PrunedConvolver::grid() const { return impl_->grid.ptr; }
bool PrunedConvolver::single() const { return impl_->single; }
size_t PrunedConvolver::bytes() const {
  return impl_->grid.count + impl_->bufA.count + impl_->bufB.count + impl_->spec.count + impl_->tw32.count +
         impl_->tw64.count;
}

fh_status PrunedConvolver::init(const int64_t N[3], const int64_t M[3], double h, bool single,
                                int max_channels, hipStream_t stream, bool &fits) {
  ConvolverImpl &c = *impl_;
  fits = false;
  c.single = single;
  c.max_channels = max_channels;
  for (int a = 0; a < 3; ++a) {
    c.N[a] = N[a];
    c.M[a] = M[a];
    c.H[a] = M[a] / 2 + 1;
    // the spectrum is computed with FP64 lines, whatever the data precision
    if (M[a] > kMaxLine || !make_plan(M[a], c.plan[a])) return FH_OK;
  }
  fits = true;
  // twiddle tables exp(-2 pi i q / M_a), host FP64, one per axis
  std::vector<double2> t64;
  for (int a = 0; a < 3; ++a) {
    c.tw_off[a] = static_cast<int64_t>(t64.size());
    for (int64_t q = 0; q < M[a]; ++q) {
      const double ang = -2.0 * 3.14159265358979323846 * static_cast<double>(q) / static_cast<double>(M[a]);
      t64.push_back(make_double2(std::cos(ang), std::sin(ang)));
    }
  }
  // digit reversal: input n = sum_s n_s (M / (r_1 ... r_s)) sits at slot
  // sum_s n_s (r_1 ... r_(s-1))
  std::vector<int> rv(t64.size());
  for (int a = 0; a < 3; ++a) {
    const LinePlan &pl = c.plan[a];
    for (int64_t n = 0; n < M[a]; ++n) {
      int64_t rest = n, weight = M[a], slot = 0, span = 1;
      for (int q = 0; q < pl.npass; ++q) {
        weight /= pl.radix[q];
        const int64_t digit = rest / weight;
        rest -= digit * weight;
        slot += digit * span;
        span *= pl.radix[q];
      }
      rv[static_cast<size_t>(c.tw_off[a] + n)] = static_cast<int>(slot);
    }
  }
  std::vector<float2> t32(t64.size());
  for (size_t q = 0; q < t64.size(); ++q)
    t32[q] = make_float2(static_cast<float>(t64[q].x), static_cast<float>(t64[q].y));
  fh_status st;
  if ((st = c.tw64.allocate(t64.size() * sizeof(double2))) || (st = c.tw32.allocate(t32.size() * sizeof(float2))))
    return st;
  if ((st = c.rev.allocate(rv.size()))) return st;
  FH_HIP_TRY(hipMemcpyAsync(c.rev.ptr, rv.data(), rv.size() * sizeof(int), hipMemcpyHostToDevice, stream));
  FH_HIP_TRY(hipMemcpyAsync(c.tw64.ptr, t64.data(), t64.size() * sizeof(double2), hipMemcpyHostToDevice, stream));
  FH_HIP_TRY(hipMemcpyAsync(c.tw32.ptr, t32.data(), t32.size() * sizeof(float2), hipMemcpyHostToDevice, stream));
  for (const void *k : {reinterpret_cast<const void *>(&line_pass_kernel<double2, true>),
                        reinterpret_cast<const void *>(&line_pass_kernel<double2, false>),
                        reinterpret_cast<const void *>(&line_pass_kernel<float2, true>),
                        reinterpret_cast<const void *>(&line_pass_kernel<float2, false>),
                        reinterpret_cast<const void *>(&column_dft_kernel<double2>),
                        reinterpret_cast<const void *>(&column_dft_kernel<float2>)})
    FH_HIP_TRY(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBudget));
  // spectrum: FP64 line passes over the full padded grid, then its even half
  {
    const int64_t Mt = M[0] * M[1] * M[2];
    DeviceArray<double2> full;
    if ((st = full.allocate(static_cast<size_t>(Mt)))) return st;
    wrapped_kernel_kernel<<<blocks_for(Mt), kThreads, 0, stream>>>(full.ptr, N[0], N[1], N[2], M[0], M[1],
                                                                   M[2], h);
    FH_HIP_TRY(hipGetLastError());
    const double2 *tw = reinterpret_cast<const double2 *>(c.tw64.ptr);
    const int m0 = static_cast<int>(M[0]), m1 = static_cast<int>(M[1]), m2 = static_cast<int>(M[2]);
    if ((st = c.pass<double2>(full.ptr, full.ptr, M[0] * M[1], 1, m2, m2, 2, 0, tw, stream)) ||
        (st = c.pass<double2>(full.ptr, full.ptr, M[0] * M[2], M[2], m1, m1, 1, 0, tw, stream)) ||
        (st = c.pass<double2>(full.ptr, full.ptr, M[1] * M[2], M[1] * M[2], m0, m0, 0, 0, tw, stream)))
      return st;
    const int64_t Ht = c.H[0] * c.H[1] * c.H[2];
    const double scale = 1.0 / static_cast<double>(Mt);
    if (single) {
      if ((st = c.spec.allocate(static_cast<size_t>(Ht) * sizeof(float)))) return st;
      half_spectrum_kernel<float><<<blocks_for(Ht), kThreads, 0, stream>>>(
          full.ptr, reinterpret_cast<float *>(c.spec.ptr), M[1], M[2], c.H[0], c.H[1], c.H[2], scale);
    } else {
      if ((st = c.spec.allocate(static_cast<size_t>(Ht) * sizeof(double)))) return st;
      half_spectrum_kernel<double><<<blocks_for(Ht), kThreads, 0, stream>>>(
          full.ptr, reinterpret_cast<double *>(c.spec.ptr), M[1], M[2], c.H[0], c.H[1], c.H[2], scale);
    }
    FH_HIP_TRY(hipGetLastError());
    FH_HIP_TRY(hipStreamSynchronize(stream));
  }
  const size_t esz = single ? sizeof(float2) : sizeof(double2);
  const size_t C = static_cast<size_t>(max_channels);
  if ((st = c.grid.allocate(C * N[0] * N[1] * N[2] * esz)) || (st = c.bufA.allocate(C * N[0] * N[1] * M[2] * esz)) ||
      (st = c.bufB.allocate(C * N[0] * M[1] * M[2] * esz)))
    return st;
  FH_HIP_TRY(hipMemsetAsync(c.bufA.ptr, 0, c.bufA.count, stream));
  FH_HIP_TRY(hipMemsetAsync(c.bufB.ptr, 0, c.bufB.count, stream));
  // tune only where the passes are long enough for the choice to matter
  if (N[0] * N[1] * M[2] >= (int64_t{1} << 20)) {
    if (single) st = c.tune<float2>(reinterpret_cast<const float2 *>(c.tw32.ptr), stream);
    else st = c.tune<double2>(reinterpret_cast<const double2 *>(c.tw64.ptr), stream);
    if (st) return st;
  }
  FH_HIP_TRY(hipStreamSynchronize(stream));
  return FH_OK;
}

void PrunedConvolver::line_groups(int32_t out[3]) const {
  out[0] = impl_->xsel[1][0];
  out[1] = impl_->xsel[0][2];
  out[2] = impl_->xsel[1][1];
}

fh_status PrunedConvolver::convolve(int channels, hipStream_t stream) {
  ConvolverImpl &c = *impl_;
  if (channels <= 0) return FH_OK;
  if (channels > c.max_channels) return set_error(FH_E_INVALID_ARGUMENT, "too many convolution channels");
  if (c.single)
    return c.convolve_t<float2>(channels, reinterpret_cast<const float2 *>(c.tw32.ptr), stream);
  return c.convolve_t<double2>(channels, reinterpret_cast<const double2 *>(c.tw64.ptr), stream);
}

}  // namespace fh
