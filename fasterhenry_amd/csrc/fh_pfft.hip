// Precorrected-FFT partial-inductance operator on gfx950: replaces
// PfftOperator (fasterhenry/src/pfft/mod.rs:270-528) with its grid
// (pfft/grid.rs), near field (pfft/near.rs) and convolution (pfft/fft.rs).
//
// y = L x is the near-field correction C x (exact partial inductance minus
// the grid's value, for pairs closer than the near-field radius) plus the
// far field mu0/4pi sum_c l_c W^T (K * W (l_c x)) through a uniform grid:
// W projects each filament's uniform current onto (order + 1)^3 Lagrange
// stencils around Gauss nodes, K(g - h) = 1/|r_g - r_h| is applied as an
// aperiodic convolution by zero-padded 3-D FFT, two real sources per complex
// transform.
//
// Set-up: the combinatorial parts (bounding box, spacing, projection boxes,
// near-pair search, CSR assembly) run on the host; the arithmetic runs on
// the GPU: projection weights (one workgroup per filament), exact near-pair
// values (the fill's pair kernels), the precorrection's grid interactions
// (one workgroup per filament row), the kernel's spectrum (rocFFT).  Every
// sum is taken in the order of the reference's definition (node order for
// the projection, source order for the box potentials), so the projection
// and the near-field corrections are bit-identical to a sequential FP64
// evaluation.
//
// Apply: near-field CSR rows (one wavefront per row, fixed-order reduction),
// spread (one thread per grid point, the transposed projection CSR read in
// filament order), FFT convolution (rocFFT, FP64), gather (one thread per
// filament).  Deterministic: no atomics on values.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "fh_fft.hpp"
#include "fh_pfft.hpp"

namespace fh {
namespace {

constexpr double kQuadTol = 1e-10;        // grid.rs:41
constexpr int kQuadMax = 12;              // grid.rs:43
constexpr size_t kMaxFftPoints = size_t{1} << 26;  // mod.rs:150
constexpr int kMaxStencil = 16;           // stencil points per axis supported here
constexpr int kNodeChunk = 128;           // projection nodes staged in LDS at a time
constexpr int kCellsPerThread = 8;
constexpr int kBoxLds = 8192;             // box potential cells held in LDS (64 KB)

// ---------------------------------------------------------------------------
// Host geometry (nalgebra operation order).

struct H3 {
  double x, y, z;
};
inline H3 h3(const double *p) { return {p[0], p[1], p[2]}; }
inline H3 operator+(H3 a, H3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline H3 operator-(H3 a, H3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline H3 operator*(H3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline H3 operator/(H3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(H3 a, H3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm(H3 a) { return std::sqrt(dot(a, a)); }
inline H3 cross(H3 a, H3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline H3 habs(H3 a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
inline double comp(H3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// neumann.rs:60-73 Bar
struct HBar {
  H3 centre, axes[3];
  double half[3];
};
HBar bar_of(const fh_filament &f) {
  HBar b;
  b.centre = (h3(f.start) + h3(f.end)) * 0.5;
  b.axes[0] = h3(f.dir);
  b.axes[1] = h3(f.wdir);
  b.axes[2] = h3(f.hdir);
  b.half[0] = 0.5 * f.length;
  b.half[1] = 0.5 * f.width;
  b.half[2] = 0.5 * f.height;
  return b;
}
double circumradius(const HBar &b) {
  return std::sqrt(b.half[0] * b.half[0] + b.half[1] * b.half[1] + b.half[2] * b.half[2]);
}
double projected_half(const HBar &b, H3 n) {
  double s = 0.0;
  for (int k = 0; k < 3; ++k) s += b.half[k] * std::fabs(dot(n, b.axes[k]));
  return s;
}
// neumann.rs:97-122: separating-axis lower bound on the gap.
double separation(const HBar &a, const HBar &b) {
  const H3 delta = b.centre - a.centre;
  const double distance = norm(delta), reach = circumradius(a) + circumradius(b);
  if (distance > 8.0 * reach) return distance - reach;
  double best = -INFINITY;
  auto consider = [&](H3 n) {
    const double gap = std::fabs(dot(n, delta)) - projected_half(a, n) - projected_half(b, n);
    best = std::fmax(best, gap);
  };
  for (int k = 0; k < 3; ++k) {
    consider(a.axes[k]);
    consider(b.axes[k]);
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const H3 n = cross(a.axes[i], b.axes[j]);
      const double nn = norm(n);
      if (nn > 1e-9) consider(n / nn);
    }
  return best;
}

// mod.rs:614-631
double auto_spacing(const double extent[3], double cells) {
  const double largest = std::fmax(std::fmax(extent[0], extent[1]), extent[2]);
  if (cells <= 1.0 || largest <= 0.0) return std::fmax(largest, 2.2250738585072014e-308);
  double small = largest / cells, large = largest;
  for (int it = 0; it < 200; ++it) {
    const double mid = std::sqrt(small * large);
    double count = 1.0;
    for (int a = 0; a < 3; ++a) count *= std::fmax(extent[a] / mid, 1.0);
    if (count > cells) small = mid;
    else large = mid;
  }
  return large;
}

// fft.rs:59-80
size_t smooth_size(size_t n) {
  size_t m = n ? n : 1;
  for (;;) {
    size_t r = m;
    while (r % 2 == 0) r /= 2;
    while (r % 3 == 0) r /= 3;
    while (r % 5 == 0) r /= 5;
    if (r == 1) return m;
    ++m;
  }
}

struct GridDesc {
  double origin[3], h;
  int64_t dims[3];
  int m;
};

// grid.rs:95-116 (host and device share the arithmetic).
__host__ __device__ inline int64_t stencil_first(const GridDesc &g, int axis, double x, double &t) {
  const double u = (x - g.origin[axis]) / g.h;
  double fl = floor(u - 0.5 * static_cast<double>(g.m)) + 1.0;
  if (fl < 0.0) fl = 0.0;
  int64_t first = static_cast<int64_t>(fl);
  if (first > g.dims[axis] - g.m) first = g.dims[axis] - g.m;
  t = u - static_cast<double>(first);
  return first;
}

__host__ __device__ inline void stencil_weights(int m, double t, double *w) {
  for (int k = 0; k < m; ++k) {
    double value = 1.0;
    for (int l = 0; l < m; ++l)
      if (l != k) value *= (t - static_cast<double>(l)) / (static_cast<double>(k) - static_cast<double>(l));
    w[k] = value;
  }
}


// Projection of one filament (grid.rs:155-216): the Gauss nodes along the
// filament's three axes, the box of grid points their stencils cover.
struct ProjMeta {
  double corner[3], axes[9], inv_area;
  int32_t count[3];   // nodes per axis
  int32_t node_off;   // into the node arrays (count[0] + count[1] + count[2] entries)
  int64_t box_lo[3];
  int32_t box_dims[3];
  int64_t out_off;    // into the dense box arrays
};

// ---------------------------------------------------------------------------
// Device kernels.

__device__ inline void node_position(const ProjMeta &pm, const double *nx, int a, int b, int c,
                                     double r[3]) {
  const double s = nx[pm.node_off + a], u = nx[pm.node_off + pm.count[0] + b],
               v = nx[pm.node_off + pm.count[0] + pm.count[1] + c];
  for (int k = 0; k < 3; ++k)
    r[k] = ((pm.corner[k] + pm.axes[k] * s) + pm.axes[3 + k] * u) + pm.axes[6 + k] * v;
}

// One workgroup per filament: W_g = sum over nodes q (in node order) of
// ((w_q lx) ly) lz on the filament's box; `touched` marks the points some
// stencil covers.
__global__ void __launch_bounds__(256) project_kernel(const ProjMeta *__restrict__ metas,
                                                      const double *__restrict__ nx,
                                                      const double *__restrict__ nw, GridDesc g,
                                                      double *__restrict__ box_w,
                                                      uint8_t *__restrict__ box_touched) {
  const ProjMeta pm = metas[blockIdx.x];
  const int m = g.m;
  const int total = pm.count[0] * pm.count[1] * pm.count[2];
  const int bx = pm.box_dims[0], by = pm.box_dims[1], bz = pm.box_dims[2];
  const int box_len = bx * by * bz;
  __shared__ double s_w[kNodeChunk];
  __shared__ int s_first[kNodeChunk][3];
  __shared__ double s_l[kNodeChunk][3][kMaxStencil];
  for (int cell_base = 0; cell_base < box_len; cell_base += 256 * kCellsPerThread) {
    double acc[kCellsPerThread];
    bool touched[kCellsPerThread];
    int cx[kCellsPerThread], cy[kCellsPerThread], cz[kCellsPerThread];
#pragma unroll
    for (int q = 0; q < kCellsPerThread; ++q) {
      acc[q] = 0.0;
      touched[q] = false;
      const int cell = cell_base + threadIdx.x + 256 * q;
      cz[q] = cell % bz;
      cy[q] = (cell / bz) % by;
      cx[q] = cell / (bz * by);
    }
    for (int node_base = 0; node_base < total; node_base += kNodeChunk) {
      __syncthreads();
      const int nn = threadIdx.x;
      if (nn < kNodeChunk && node_base + nn < total) {
        const int idx = node_base + nn;
        const int c = idx % pm.count[2], b = (idx / pm.count[2]) % pm.count[1],
                  a = idx / (pm.count[2] * pm.count[1]);
        double r[3];
        node_position(pm, nx, a, b, c, r);
        s_w[nn] = nw[pm.node_off + a] * nw[pm.node_off + pm.count[0] + b] *
                  nw[pm.node_off + pm.count[0] + pm.count[1] + c] * pm.inv_area;
        for (int ax = 0; ax < 3; ++ax) {
          double t;
          s_first[nn][ax] = static_cast<int>(stencil_first(g, ax, r[ax], t) - pm.box_lo[ax]);
          stencil_weights(m, t, s_l[nn][ax]);
        }
      }
      __syncthreads();
      const int chunk = min(kNodeChunk, total - node_base);
      for (int k = 0; k < chunk; ++k) {
        const int fx = s_first[k][0], fy = s_first[k][1], fz = s_first[k][2];
        const double w = s_w[k];
#pragma unroll
        for (int q = 0; q < kCellsPerThread; ++q) {
          const int dx = cx[q] - fx, dy = cy[q] - fy, dz = cz[q] - fz;
          if (dx >= 0 && dx < m && dy >= 0 && dy < m && dz >= 0 && dz < m) {
            acc[q] += ((w * s_l[k][0][dx]) * s_l[k][1][dy]) * s_l[k][2][dz];
            touched[q] = true;
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < kCellsPerThread; ++q) {
      const int cell = cell_base + threadIdx.x + 256 * q;
      if (cell < box_len) {
        box_w[pm.out_off + cell] = acc[q];
        box_touched[pm.out_off + cell] = touched[q] ? 1 : 0;
      }
    }
  }
}

// Precorrection grid interactions (near.rs:316-389) of one row: for each
// partner j, sum_h W_jh sum_g W_ig K(h - g), by the direct double sum or
// through the potential of i's projection over the partners' box --
// whichever the reference's cost rule picks -- in its order of addition.
struct RowTask {
  int64_t row;          // filament i
  int64_t partner_off;  // into partner lists / approx outputs
  int32_t partners;
  int32_t use_box;
  int64_t box_lo[3];
  int32_t box_dims[3];
  int64_t scratch_off;  // psi in the scratch array (box rows too big for LDS)
};

__device__ inline void grid_point(int64_t gidx, const GridDesc &g, int64_t p[3]) {
  p[2] = gidx % g.dims[2];
  const int64_t rest = gidx / g.dims[2];
  p[1] = rest % g.dims[1];
  p[0] = rest / g.dims[1];
}

// K(d) for every offset a row's evaluation can meet, |d_a| <= r_a, computed
// by pfft_grid_kernel itself (so a lookup equals the direct evaluation).
struct KernelTable {
  const double *t;
  int32_t r0, r1, r2, d1, d2;
};

__global__ void kernel_table_kernel(double *__restrict__ t, int32_t r0, int32_t r1, int32_t r2, double h) {
  const int64_t d0 = 2 * r0 + 1, d1 = 2 * r1 + 1, d2 = 2 * r2 + 1;
  const int64_t total = d0 * d1 * d2;
  for (int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; q < total;
       q += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t z = q % d2, y = (q / d2) % d1, x = q / (d2 * d1);
    t[q] = pfft_grid_kernel(x - r0, y - r1, z - r2, h);
  }
}

template <bool TABLE>
__device__ inline double kval(const KernelTable &T, const GridDesc &g, int dx, int dy, int dz) {
  if (TABLE) return T.t[(static_cast<int64_t>(dx + T.r0) * T.d1 + (dy + T.r1)) * T.d2 + (dz + T.r2)];
  return pfft_grid_kernel(dx, dy, dz, g.h);
}

constexpr int kSrcLds = 1024;  // source entries of a row staged in LDS

// Row i's projection entries as grid coordinates and weights, in entry
// order, in LDS (rows with more than kSrcLds entries read them from memory).
struct SrcStage {
  int x[kSrcLds], y[kSrcLds], z[kSrcLds];
  double w[kSrcLds];
};

template <bool STAGED>
__device__ inline void source(const SrcStage &S, int64_t s0, int s, const uint32_t *proj_idx,
                              const double *proj_w, const GridDesc &g, int &x, int &y, int &z, double &w) {
  if (STAGED) {
    x = S.x[s];
    y = S.y[s];
    z = S.z[s];
    w = S.w[s];
  } else {
    int64_t p[3];
    grid_point(proj_idx[s0 + s], g, p);
    x = static_cast<int>(p[0]);
    y = static_cast<int>(p[1]);
    z = static_cast<int>(p[2]);
    w = proj_w[s0 + s];
  }
}

__device__ inline void stage_sources(SrcStage &S, int64_t s0, int ns, const uint32_t *proj_idx,
                                     const double *proj_w, const GridDesc &g) {
  for (int s = threadIdx.x; s < ns && s < kSrcLds; s += blockDim.x) {
    int64_t p[3];
    grid_point(proj_idx[s0 + s], g, p);
    S.x[s] = static_cast<int>(p[0]);
    S.y[s] = static_cast<int>(p[1]);
    S.z[s] = static_cast<int>(p[2]);
    S.w[s] = proj_w[s0 + s];
  }
  __syncthreads();
}

// Direct rows: one thread per partner, the double sum in the reference's
// order (partner entries outer, source entries inner).
template <bool STAGED, bool TABLE>
__device__ void direct_row(const RowTask &t, const SrcStage &S, int64_t s0, int ns,
                           const int64_t *__restrict__ proj_off, const uint32_t *__restrict__ proj_idx,
                           const double *__restrict__ proj_w, const int64_t *__restrict__ partner_idx,
                           const GridDesc &g, const KernelTable &T, double *__restrict__ approx) {
  for (int q = threadIdx.x; q < t.partners; q += blockDim.x) {
    const int64_t j = partner_idx[t.partner_off + q];
    double total = 0.0;
    for (int64_t e = proj_off[j]; e < proj_off[j + 1]; ++e) {
      int64_t hp[3];
      grid_point(proj_idx[e], g, hp);
      double potential = 0.0;
      for (int s = 0; s < ns; ++s) {
        int x, y, z;
        double w;
        source<STAGED>(S, s0, s, proj_idx, proj_w, g, x, y, z, w);
        potential += w * kval<TABLE>(T, g, static_cast<int>(hp[0]) - x, static_cast<int>(hp[1]) - y,
                                     static_cast<int>(hp[2]) - z);
      }
      total += proj_w[e] * potential;
    }
    approx[t.partner_off + q] = total;
  }
}

template <bool TABLE>
__global__ void __launch_bounds__(256) interactions_direct_kernel(
    const RowTask *__restrict__ tasks, const int64_t *__restrict__ proj_off,
    const uint32_t *__restrict__ proj_idx, const double *__restrict__ proj_w,
    const int64_t *__restrict__ partner_idx, GridDesc g, KernelTable T, double *__restrict__ approx) {
  __shared__ SrcStage S;
  const RowTask t = tasks[blockIdx.x];
  const int64_t s0 = proj_off[t.row];
  const int ns = static_cast<int>(proj_off[t.row + 1] - s0);
  stage_sources(S, s0, ns, proj_idx, proj_w, g);
  if (ns <= kSrcLds) direct_row<true, TABLE>(t, S, s0, ns, proj_off, proj_idx, proj_w, partner_idx, g, T, approx);
  else direct_row<false, TABLE>(t, S, s0, ns, proj_off, proj_idx, proj_w, partner_idx, g, T, approx);
}

// Box rows: the potential psi of row i's projection over the box of its
// partners' entries (each cell sums the sources in order; fused multiply-add
// for the first bz - bz % 4 cells of a z-row, as near.rs:189-207 does in
// blocks of four, plain multiply-add for the tail), held in LDS when it fits
// (`lds_cells`), else in `scratch`; then one thread per partner.
template <bool STAGED, bool TABLE>
__device__ void box_potential(const RowTask &t, const SrcStage &S, int64_t s0, int ns,
                              const uint32_t *__restrict__ proj_idx, const double *__restrict__ proj_w,
                              const GridDesc &g, const KernelTable &T, double *psi) {
  const int bx = t.box_dims[0], by = t.box_dims[1], bz = t.box_dims[2];
  const int box_len = bx * by * bz;
  const int body = bz - bz % 4;
  for (int cell = threadIdx.x; cell < box_len; cell += blockDim.x) {
    const int z = cell % bz, y = (cell / bz) % by, x = cell / (bz * by);
    const int px = static_cast<int>(t.box_lo[0]) + x, py = static_cast<int>(t.box_lo[1]) + y,
              pz = static_cast<int>(t.box_lo[2]) + z;
    double acc = 0.0;
    if (z < body) {
      for (int s = 0; s < ns; ++s) {
        int sx, sy, sz;
        double w;
        source<STAGED>(S, s0, s, proj_idx, proj_w, g, sx, sy, sz, w);
        acc = __builtin_fma(w, kval<TABLE>(T, g, px - sx, py - sy, pz - sz), acc);
      }
    } else {
      for (int s = 0; s < ns; ++s) {
        int sx, sy, sz;
        double w;
        source<STAGED>(S, s0, s, proj_idx, proj_w, g, sx, sy, sz, w);
        acc += w * kval<TABLE>(T, g, px - sx, py - sy, pz - sz);
      }
    }
    psi[cell] = acc;
  }
}

template <bool TABLE>
__global__ void __launch_bounds__(512) interactions_box_kernel(
    const RowTask *__restrict__ tasks, const int64_t *__restrict__ proj_off,
    const uint32_t *__restrict__ proj_idx, const double *__restrict__ proj_w,
    const int64_t *__restrict__ partner_idx, GridDesc g, KernelTable T, int lds_cells,
    double *__restrict__ scratch, double *__restrict__ approx) {
  extern __shared__ double psi_lds[];
  __shared__ SrcStage S;
  const RowTask t = tasks[blockIdx.x];
  const int64_t s0 = proj_off[t.row];
  const int ns = static_cast<int>(proj_off[t.row + 1] - s0);
  const int by = t.box_dims[1], bz = t.box_dims[2];
  const int box_len = t.box_dims[0] * by * bz;
  double *psi = box_len <= lds_cells ? psi_lds : scratch + t.scratch_off;
  stage_sources(S, s0, ns, proj_idx, proj_w, g);
  if (ns <= kSrcLds) box_potential<true, TABLE>(t, S, s0, ns, proj_idx, proj_w, g, T, psi);
  else box_potential<false, TABLE>(t, S, s0, ns, proj_idx, proj_w, g, T, psi);
  __syncthreads();
  for (int q = threadIdx.x; q < t.partners; q += blockDim.x) {
    const int64_t j = partner_idx[t.partner_off + q];
    double total = 0.0;
    for (int64_t e = proj_off[j]; e < proj_off[j + 1]; ++e) {
      int64_t p[3];
      grid_point(proj_idx[e], g, p);
      const int64_t local =
          ((p[0] - t.box_lo[0]) * by + (p[1] - t.box_lo[1])) * bz + (p[2] - t.box_lo[2]);
      total += proj_w[e] * psi[local];
    }
    approx[t.partner_off + q] = total;
  }
}

// The kernel on the padded grid with wrapped offsets (fft.rs:101-128).
__global__ void kernel_grid_kernel(double2 *__restrict__ data, int64_t n0, int64_t n1, int64_t n2,
                                   int64_t m0, int64_t m1, int64_t m2, double h) {
  const int64_t total = m0 * m1 * m2;
  for (int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; q < total;
       q += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t k = q % m2, j = (q / m2) % m1, i = q / (m2 * m1);
    auto off = [](int64_t x, int64_t n, int64_t m, bool &ok) -> int64_t {
      ok = true;
      if (x < n) return x;
      if (x + n > m) return x - m;
      ok = false;
      return 0;
    };
    bool a, b, c;
    const int64_t di = off(i, n0, m0, a), dj = off(j, n1, m1, b), dk = off(k, n2, m2, c);
    data[q] = make_double2(a && b && c ? pfft_grid_kernel(di, dj, dk, h) : 0.0, 0.0);
  }
}

__global__ void spectrum_kernel(const double2 *__restrict__ data, double *__restrict__ spectrum,
                                int64_t total, double scale) {
  for (int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; q < total;
       q += static_cast<int64_t>(gridDim.x) * blockDim.x)
    spectrum[q] = data[q].x * scale;
}

// Near field: y_v[i] = sum_e C_e x_v[col_e], one wavefront per row, lanes
// over the entries, partial sums added in a fixed tree.
__global__ void __launch_bounds__(256) near_kernel(const int64_t *__restrict__ off,
                                                   const uint32_t *__restrict__ col,
                                                   const double *__restrict__ val, int64_t n,
                                                   int V, const double *__restrict__ x,
                                                   double *__restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int64_t row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= n) return;
  const int64_t e0 = off[row], e1 = off[row + 1];
  for (int v = 0; v < V; ++v) {
    const double *xv = x + static_cast<int64_t>(v) * n;
    double s = 0.0;
    for (int64_t e = e0 + lane; e < e1; e += 64) s += val[e] * xv[col[e]];
    for (int d = 32; d > 0; d >>= 1) s += __shfl_xor(s, d);
    if (lane == 0) y[static_cast<int64_t>(v) * n + row] = s;
  }
}

// Sources into the padded FFT buffer: re = q_a, im = q_b (zero outside the
// grid), q(g) = sum over the transposed projection (filament order) of
// (w l_c) x_v.
__global__ void spread_kernel(double2 *__restrict__ data, const int64_t *__restrict__ soff,
                              const uint32_t *__restrict__ sfil, const double *__restrict__ sw,
                              const double *__restrict__ dirs, const double *__restrict__ x,
                              int64_t n, int64_t n0, int64_t n1, int64_t n2, int64_t m1,
                              int64_t m2, int64_t total, int va, int ca, int vb, int cb) {
  for (int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; q < total;
       q += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t k = q % m2, j = (q / m2) % m1, i = q / (m2 * m1);
    double a = 0.0, b = 0.0;
    if (i < n0 && j < n1 && k < n2) {
      const int64_t gidx = (i * n1 + j) * n2 + k;
      const double *xa = x + static_cast<int64_t>(va) * n;
      const double *xb = x + static_cast<int64_t>(vb >= 0 ? vb : 0) * n;
      for (int64_t e = soff[gidx]; e < soff[gidx + 1]; ++e) {
        const int64_t f = sfil[e];
        a += sw[e] * dirs[3 * f + ca] * xa[f];
        if (vb >= 0) b += sw[e] * dirs[3 * f + cb] * xb[f];
      }
    }
    data[q] = make_double2(a, b);
  }
}

__global__ void multiply_kernel(double2 *__restrict__ data, const double *__restrict__ spectrum,
                                int64_t total) {
  for (int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; q < total;
       q += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const double s = spectrum[q];
    double2 d = data[q];
    d.x *= s;
    d.y *= s;
    data[q] = d;
  }
}

// far_v[i] += l_c(i) sum_g W_ig phi(g) for the one or two sources of a pass.
__global__ void gather_kernel(const double2 *__restrict__ data, const int64_t *__restrict__ poff,
                              const uint32_t *__restrict__ pidx, const double *__restrict__ pw,
                              const double *__restrict__ dirs, int64_t n, int64_t n1, int64_t n2,
                              int64_t m1, int64_t m2, double *__restrict__ far, int va, int ca,
                              int vb, int cb) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  double sa = 0.0, sb = 0.0;
  for (int64_t e = poff[i]; e < poff[i + 1]; ++e) {
    const int64_t gidx = pidx[e];
    const int64_t k = gidx % n2, j = (gidx / n2) % n1, ii = gidx / (n2 * n1);
    const double2 phi = data[(ii * m1 + j) * m2 + k];
    sa += pw[e] * phi.x;
    if (vb >= 0) sb += pw[e] * phi.y;
  }
  far[static_cast<int64_t>(va) * n + i] += dirs[3 * i + ca] * sa;
  if (vb >= 0) far[static_cast<int64_t>(vb) * n + i] += dirs[3 * i + cb] * sb;
}

__global__ void finish_kernel(double *__restrict__ y, const double *__restrict__ far, int64_t count) {
  const int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (q < count) y[q] += 1.0e-7 * far[q];
}

__global__ void zero_kernel(double *__restrict__ p, int64_t count) {
  const int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (q < count) p[q] = 0.0;
}

// Up to 8 sources (vector, direction component) of one convolution batch,
// paired into channels (2c, 2c + 1).
constexpr int kBatchSources = 8;
struct SourceSet {
  int count;
  int vec[kBatchSources];
  int comp[kBatchSources];
};

__device__ inline float2 to_c(double a, double b, float2) { return make_float2(static_cast<float>(a), static_cast<float>(b)); }
__device__ inline double2 to_c(double a, double b, double2) { return make_double2(a, b); }

// The batch's per-filament source factors packed together: pk[f][s] =
// (l_c(f), x_v(f)) for source s = (v, c), so that the spread reads one
// contiguous record per projection entry instead of gathering the direction
// and every vector separately.
__global__ void pack_sources_kernel(const double *__restrict__ dirs, const double *__restrict__ x,
                                    int64_t n, SourceSet ss, double2 *__restrict__ pk) {
  const int64_t f = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (f >= n) return;
#pragma unroll
  for (int s = 0; s < kBatchSources; ++s)
    if (s < ss.count) pk[f * ss.count + s] = make_double2(dirs[3 * f + ss.comp[s]], x[ss.vec[s] * n + f]);
}

// Sources on the compact grid: q_s(g) = sum over the transposed projection
// (filaments in order) of (w l_c) x_v, in FP64, stored as channel values.
// Workgroup b runs on XCD b mod 8; it takes the block of grid points
// (b mod 8) * per + b / 8 (the grid is a multiple of 8 workgroups), so each
// XCD spreads one contiguous eighth of the grid and the packed records of
// its filaments stay in that XCD's L2 while the CSR streams through.
template <typename T2>
__global__ void __launch_bounds__(256) spread_compact_kernel(
    T2 *__restrict__ grid, const int64_t *__restrict__ soff, const uint32_t *__restrict__ sfil,
    const double *__restrict__ sw, const double2 *__restrict__ pk, int64_t G, SourceSet ss) {
  const int64_t per = gridDim.x / 8;
  const int64_t block = static_cast<int64_t>(blockIdx.x % 8) * per + blockIdx.x / 8;
  const int64_t g = block * blockDim.x + threadIdx.x;
  if (g >= G) return;
  double acc[kBatchSources];
#pragma unroll
  for (int s = 0; s < kBatchSources; ++s) acc[s] = 0.0;
  for (int64_t e = soff[g]; e < soff[g + 1]; ++e) {
    const double2 *q = pk + static_cast<int64_t>(sfil[e]) * ss.count;
    const double w = sw[e];
#pragma unroll
    for (int s = 0; s < kBatchSources; ++s)
      if (s < ss.count) {
        const double2 dx = q[s];
        acc[s] += w * dx.x * dx.y;
      }
  }
#pragma unroll
  for (int c = 0; c < kBatchSources / 2; ++c)
    if (2 * c < ss.count) grid[c * G + g] = to_c(acc[2 * c], acc[2 * c + 1], T2{});
}

// far_v[i] += l_c(i) sum_e W_e phi_s(g_e) for the batch's sources (the
// components of one vector in ascending order): one wavefront per filament,
// lanes striding its projection entries, partial sums combined in a fixed
// butterfly order.  Filaments are dealt to XCDs in contiguous eighths (as
// in the spread), so neighbouring filaments read their grid points from
// one L2.
template <typename T2>
__global__ void __launch_bounds__(256) gather_compact_kernel(
    const T2 *__restrict__ grid, const int64_t *__restrict__ poff, const uint32_t *__restrict__ pidx,
    const double *__restrict__ pw, const double *__restrict__ dirs, int64_t n, int64_t G, SourceSet ss,
    double *__restrict__ far) {
  const int lane = threadIdx.x & 63;
  const int64_t per = gridDim.x / 8;
  const int64_t block = static_cast<int64_t>(blockIdx.x % 8) * per + blockIdx.x / 8;
  const int64_t i = block * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  double sum[kBatchSources];
#pragma unroll
  for (int s = 0; s < kBatchSources; ++s) sum[s] = 0.0;
  const int C = (ss.count + 1) / 2;
  for (int64_t e = poff[i] + lane; e < poff[i + 1]; e += 64) {
    const int64_t g = pidx[e];
    const double w = pw[e];
#pragma unroll
    for (int c = 0; c < kBatchSources / 2; ++c)
      if (c < C) {
        const T2 v = grid[c * G + g];
        sum[2 * c] += w * static_cast<double>(v.x);
        sum[2 * c + 1] += w * static_cast<double>(v.y);
      }
  }
#pragma unroll
  for (int s = 0; s < kBatchSources; ++s)
    if (s < ss.count)
      for (int d = 32; d > 0; d >>= 1) sum[s] += __shfl_xor(sum[s], d);
  if (lane == 0) {
#pragma unroll
    for (int s = 0; s < kBatchSources; ++s)
      if (s < ss.count) far[ss.vec[s] * n + i] += dirs[3 * i + ss.comp[s]] * sum[s];
  }
}

unsigned blocks_for(int64_t total) {
  return static_cast<unsigned>(std::min<int64_t>((total + 255) / 256, 1 << 20));
}
// Workgroups of 256 covering `total` threads, rounded up to a multiple of 8
// (one share per XCD; the kernel maps them back to contiguous blocks).
unsigned xcd_blocks_for(int64_t total) { return static_cast<unsigned>(((total + 255) / 256 + 7) / 8 * 8); }

// Host thread pool helper: f(begin, end) over [0, n).
template <typename F>
void parallel_for(size_t n, F f) {
  unsigned threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 1024 || threads == 1) {
    f(size_t{0}, n);
    return;
  }
  std::vector<std::thread> pool;
  std::atomic<size_t> next{0};
  const size_t grain = std::max<size_t>(64, n / (threads * 16));
  for (unsigned t = 0; t < threads; ++t)
    pool.emplace_back([&]() {
      for (;;) {
        const size_t b = next.fetch_add(grain);
        if (b >= n) break;
        f(b, std::min(n, b + grain));
      }
    });
  for (auto &th : pool) th.join();
}

// ---------------------------------------------------------------------------
// The FP64 3-D FFT of the padded grid for grids with an axis too long for
// one LDS line (more than 8 191 points): Stockham autosort passes through
// global memory, one launch per radix (4, 2, 3, 5) and axis, ping-ponging
// between the grid and a work buffer.  Slower than the LDS line passes, and
// only used for such extreme aspect ratios.

// v[0..R) <- DFT_R with kernel exp(sgn 2 pi i n k / R).
template <int R>
__device__ inline void small_dft(double2 *v, double sgn) {
  if constexpr (R == 2) {
    const double2 a = v[0], b = v[1];
    v[0] = make_double2(a.x + b.x, a.y + b.y);
    v[1] = make_double2(a.x - b.x, a.y - b.y);
  } else if constexpr (R == 4) {
    const double2 t0 = make_double2(v[0].x + v[2].x, v[0].y + v[2].y);
    const double2 t1 = make_double2(v[0].x - v[2].x, v[0].y - v[2].y);
    const double2 t2 = make_double2(v[1].x + v[3].x, v[1].y + v[3].y);
    const double2 d = make_double2(v[1].x - v[3].x, v[1].y - v[3].y);
    const double2 wd = make_double2(-sgn * d.y, sgn * d.x);  // (sgn i) d
    v[0] = make_double2(t0.x + t2.x, t0.y + t2.y);
    v[2] = make_double2(t0.x - t2.x, t0.y - t2.y);
    v[1] = make_double2(t1.x + wd.x, t1.y + wd.y);
    v[3] = make_double2(t1.x - wd.x, t1.y - wd.y);
  } else if constexpr (R == 3) {
    const double h = 0.86602540378443864676;  // sqrt(3) / 2
    const double2 t = make_double2(v[1].x + v[2].x, v[1].y + v[2].y);
    const double2 u = make_double2(v[1].x - v[2].x, v[1].y - v[2].y);
    const double2 m = make_double2(v[0].x - 0.5 * t.x, v[0].y - 0.5 * t.y);
    const double2 iu = make_double2(-sgn * h * u.y, sgn * h * u.x);  // sgn i (sqrt3/2) u
    v[0] = make_double2(v[0].x + t.x, v[0].y + t.y);
    v[1] = make_double2(m.x + iu.x, m.y + iu.y);
    v[2] = make_double2(m.x - iu.x, m.y - iu.y);
  } else {  // R == 5
    const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;  // cos 2pi/5, cos 4pi/5
    const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;   // sin 2pi/5, sin 4pi/5
    const double2 a1 = make_double2(v[1].x + v[4].x, v[1].y + v[4].y);
    const double2 b1 = make_double2(v[1].x - v[4].x, v[1].y - v[4].y);
    const double2 a2 = make_double2(v[2].x + v[3].x, v[2].y + v[3].y);
    const double2 b2 = make_double2(v[2].x - v[3].x, v[2].y - v[3].y);
    const double2 x0 = v[0];
    const double2 p1 = make_double2(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
    const double2 p2 = make_double2(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
    // sgn i (s1 b1 + s2 b2) and sgn i (s2 b1 - s1 b2)
    const double2 q1 = make_double2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y);
    const double2 q2 = make_double2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y);
    const double2 iq1 = make_double2(-sgn * q1.y, sgn * q1.x), iq2 = make_double2(-sgn * q2.y, sgn * q2.x);
    v[0] = make_double2(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
    v[1] = make_double2(p1.x + iq1.x, p1.y + iq1.y);
    v[4] = make_double2(p1.x - iq1.x, p1.y - iq1.y);
    v[2] = make_double2(p2.x + iq2.x, p2.y + iq2.y);
    v[3] = make_double2(p2.x - iq2.x, p2.y - iq2.y);
  }
}

// One radix-R Stockham pass along an axis of length M and element stride s
// (lines = total / M): y[d + r Ns] = DFT_R(x[j + r M / R] w^(k r)), k = j mod Ns,
// d = (j / Ns) Ns R + k.  Consecutive threads take consecutive inner offsets,
// so both the strided axes and the contiguous one load coalesced.
template <int R>
__global__ void stockham_pass_kernel(const double2 *__restrict__ x, double2 *__restrict__ y, int64_t total,
                                     int64_t M, int64_t s, int64_t Ns, const double2 *__restrict__ tw,
                                     double sgn) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  const int64_t mr = M / R;
  if (t >= total / R) return;
  const int64_t inner = t % s, rest = t / s;
  const int64_t j = rest % mr, outer = rest / mr;
  const int64_t base = outer * M * s + inner;
  const int64_t k = j % Ns;
  const int64_t step = M / (Ns * R);
  double2 v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    double2 a = x[base + (j + r * mr) * s];
    if (r > 0 && k > 0) {
      const double2 w = tw[k * r * step];
      const double wy = sgn < 0.0 ? w.y : -w.y;  // the table holds exp(-2 pi i q / M)
      a = make_double2(a.x * w.x - a.y * wy, a.x * wy + a.y * w.x);
    }
    v[r] = a;
  }
  small_dft<R>(v, sgn);
  const int64_t d = (j / Ns) * Ns * R + k;
#pragma unroll
  for (int r = 0; r < R; ++r) y[base + (d + r * Ns) * s] = v[r];
}

struct StockhamAxis {
  int64_t M = 0;
  std::vector<int> radix;
  DeviceArray<double2> tw;  // exp(-2 pi i q / M), q < M
};

fh_status stockham_axis_plan(int64_t M, StockhamAxis &ax, hipStream_t stream) {
  ax.M = M;
  ax.radix.clear();
  int64_t r = M;
  for (const int f : {4, 2, 3, 5})
    while (r % f == 0) {
      ax.radix.push_back(f);
      r /= f;
    }
  if (r != 1) return set_error(FH_E_PFFT, "padded FFT length " + std::to_string(M) + " is not 2-3-5 smooth");
  std::vector<double2> h(static_cast<size_t>(M));
  for (int64_t q = 0; q < M; ++q) {
    const double a = -6.283185307179586476925 * static_cast<double>(q) / static_cast<double>(M);
    h[q] = make_double2(std::cos(a), std::sin(a));
  }
  fh_status st;
  if ((st = ax.tw.allocate(static_cast<size_t>(M)))) return st;
  FH_HIP_TRY(hipMemcpyAsync(ax.tw.ptr, h.data(), M * sizeof(double2), hipMemcpyHostToDevice, stream));
  FH_HIP_TRY(hipStreamSynchronize(stream));  // h is read before it goes out of scope
  return FH_OK;
}

// The unnormalized 3-D DFT (kernel exp(-+2 pi i k.n / M)) of data
// [M0][M1][M2], in place (work: the same size).
fh_status stockham_3d(double2 *data, double2 *work, const int64_t dims[3], StockhamAxis ax[3],
                      bool forward, hipStream_t stream) {
  const int64_t total = dims[0] * dims[1] * dims[2];
  const double sgn = forward ? -1.0 : 1.0;
  double2 *src = data, *dst = work;
  for (int a = 2; a >= 0; --a) {
    const int64_t M = dims[a];
    const int64_t s = a == 2 ? 1 : (a == 1 ? dims[2] : dims[1] * dims[2]);
    int64_t Ns = 1;
    for (const int R : ax[a].radix) {
      const int64_t threads = total / R;
      const unsigned blocks = static_cast<unsigned>((threads + 255) / 256);
      if (R == 4)
        stockham_pass_kernel<4><<<blocks, 256, 0, stream>>>(src, dst, total, M, s, Ns, ax[a].tw.ptr, sgn);
      else if (R == 2)
        stockham_pass_kernel<2><<<blocks, 256, 0, stream>>>(src, dst, total, M, s, Ns, ax[a].tw.ptr, sgn);
      else if (R == 3)
        stockham_pass_kernel<3><<<blocks, 256, 0, stream>>>(src, dst, total, M, s, Ns, ax[a].tw.ptr, sgn);
      else
        stockham_pass_kernel<5><<<blocks, 256, 0, stream>>>(src, dst, total, M, s, Ns, ax[a].tw.ptr, sgn);
      FH_HIP_TRY(hipGetLastError());
      Ns *= R;
      std::swap(src, dst);
    }
  }
  if (src != data)
    FH_HIP_TRY(hipMemcpyAsync(data, src, total * sizeof(double2), hipMemcpyDeviceToDevice, stream));
  return FH_OK;
}


}  // namespace

struct PfftImpl {
  size_t n = 0;
  bool has_grid = false;
  GridDesc g{};
  int64_t padded[3] = {0, 0, 0};
  int comps[3] = {0, 0, 0}, ncomp = 0;
  size_t proj_entries = 0, near_entries = 0;
  DeviceArray<double> dirs;
  DeviceArray<int64_t> proj_off, spread_off, near_off;
  DeviceArray<uint32_t> proj_idx, spread_fil, near_col;
  DeviceArray<double> proj_w, spread_w, near_val;
  DeviceArray<double> spectrum;
  DeviceArray<double2> buffer;
  DeviceArray<double> far;
  DeviceArray<double2> packed;  // n x kBatchSources source factors
  DeviceArray<double2> work;
  StockhamAxis axes[3];
  size_t far_cap = 0;
  PrunedConvolver conv;
  bool use_conv = false;
  int fft_kind = 0;

  int64_t padded_total() const { return padded[0] * padded[1] * padded[2]; }
  int64_t grid_total() const { return g.dims[0] * g.dims[1] * g.dims[2]; }

  fh_status fft(bool forward, hipStream_t stream) {
    return stockham_3d(buffer.ptr, work.ptr, padded, axes, forward, stream);
  }

};

PfftOperator::PfftOperator() : impl_(new PfftImpl) {}
PfftOperator::~PfftOperator() = default;
size_t PfftOperator::size() const { return impl_->n; }

fh_status PfftOperator::build(const std::vector<fh_filament> &F, const fh_pfft_params &prm,
                              hipStream_t stream) {
  PfftImpl &op = *impl_;
  const int order = prm.interpolation_order;
  if (order < 1)
    return set_error(FH_E_PFFT, "invalid pFFT parameter: interpolation_order must be at least 1",
                     0, 0, 0.0, 1);
  if (!(std::isfinite(prm.near_field_radius) && prm.near_field_radius >= order + 1.0))
    return set_error(FH_E_PFFT,
                     "invalid pFFT parameter: near_field_radius must be finite and at least "
                     "interpolation_order + 1",
                     0, 0, 0.0, 1);
  if (prm.auto_grid) {
    if (!(prm.cells_per_filament > 0.0 && std::isfinite(prm.cells_per_filament)))
      return set_error(FH_E_PFFT, "invalid pFFT parameter: cells_per_filament must be positive and finite", 0, 0, 0.0, 1);
  } else if (!(prm.spacing > 0.0 && std::isfinite(prm.spacing))) {
    return set_error(FH_E_PFFT, "invalid pFFT parameter: grid spacing must be positive and finite", 0, 0, 0.0, 1);
  }
  if (order + 1 > kMaxStencil)
    return set_error(FH_E_UNSUPPORTED, "interpolation orders above 15 are not provided");
  const size_t n = F.size();
  op.n = n;
  fh_status st;
  if ((st = op.proj_off.allocate(n + 1)) || (st = op.near_off.allocate(n + 1))) return st;
  if (n == 0) {
    FH_HIP_TRY(hipMemsetAsync(op.proj_off.ptr, 0, sizeof(int64_t), stream));
    FH_HIP_TRY(hipMemsetAsync(op.near_off.ptr, 0, sizeof(int64_t), stream));
    FH_HIP_TRY(hipStreamSynchronize(stream));
    return FH_OK;
  }
  // mod.rs:580-593 bounding box
  H3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
  for (const auto &f : F) {
    const H3 reach = (habs(h3(f.wdir)) * f.width + habs(h3(f.hdir)) * f.height) * 0.5;
    for (const H3 e : {h3(f.start), h3(f.end)}) {
      const H3 a = e - reach, b = e + reach;
      lo = {std::fmin(lo.x, a.x), std::fmin(lo.y, a.y), std::fmin(lo.z, a.z)};
      hi = {std::fmax(hi.x, b.x), std::fmax(hi.y, b.y), std::fmax(hi.z, b.z)};
    }
  }
  const double ext[3] = {hi.x - lo.x, hi.y - lo.y, hi.z - lo.z};
  const double h = prm.auto_grid ? auto_spacing(ext, prm.cells_per_filament * static_cast<double>(n))
                                 : prm.spacing;
  // grid.rs:64-80
  GridDesc &g = op.g;
  g.m = order + 1;
  g.h = h;
  const double pad = static_cast<double>((g.m + 1) / 2);
  const double lo3[3] = {lo.x, lo.y, lo.z}, hi3[3] = {hi.x, hi.y, hi.z};
  size_t padded_total = 1, grid_total = 1;
  for (int a = 0; a < 3; ++a) {
    g.origin[a] = lo3[a] - pad * h;
    const double span = (hi3[a] - g.origin[a]) / h;
    g.dims[a] = static_cast<int64_t>(std::floor(span + 0.5 * g.m)) + 2;
    op.padded[a] = static_cast<int64_t>(smooth_size(2 * static_cast<size_t>(g.dims[a]) - 1));
    padded_total *= static_cast<size_t>(op.padded[a]);
    grid_total *= static_cast<size_t>(g.dims[a]);
  }
  if (padded_total > kMaxFftPoints || grid_total > 0xffffffffull)
    return set_error(FH_E_PFFT,
                     "pFFT grid of " + std::to_string(padded_total) +
                         " padded points exceeds the limit of " + std::to_string(kMaxFftPoints),
                     static_cast<int64_t>(padded_total), static_cast<int64_t>(kMaxFftPoints), 0.0, 2);
  op.has_grid = true;
  const double radius = prm.near_field_radius * h;

  // ---- projection: nodes (host), boxes (host), weights (device) ----
  double thresholds[kQuadMax];
  for (int k = 1; k <= kQuadMax; ++k) thresholds[k - 1] = std::sinh(std::log(1.0 / kQuadTol) / (2.0 * k));
  auto order_for = [&](double half, double distance) {
    if (half <= 0.0) return 1;
    if (distance != distance || distance <= 0.0) return 0;
    const double delta = distance / half;
    for (int i = 0; i < kQuadMax; ++i)
      if (delta >= thresholds[i]) return i + 1;
    return 0;
  };
  double rule_x[kQuadMax + 1][kQuadMax], rule_w[kQuadMax + 1][kQuadMax];
  for (int k = 1; k <= kQuadMax; ++k) gauss_rule_host(k, rule_x[k], rule_w[k]);
  std::vector<ProjMeta> metas(n);
  std::vector<int32_t> node_counts(n);
  for (size_t i = 0; i < n; ++i) {
    const fh_filament &f = F[i];
    const double extents[3] = {f.length, f.width, f.height};
    int32_t c = 0;
    for (int d = 0; d < 3; ++d) {
      double pc = std::ceil(extents[d] / h);
      if (pc < 1.0) pc = 1.0;
      int o = order_for(0.5 * (extents[d] / pc), radius);
      if (!o) o = kQuadMax;
      metas[i].count[d] = static_cast<int32_t>(pc) * o;
      c += metas[i].count[d];
    }
    node_counts[i] = c;
  }
  std::vector<int64_t> node_off(n + 1, 0);
  for (size_t i = 0; i < n; ++i) node_off[i + 1] = node_off[i] + node_counts[i];
  if (node_off[n] > 0x7fffffff) return set_error(FH_E_PFFT, "pFFT projection too large", 0, 0, 0.0, 2);
  std::vector<double> nx(static_cast<size_t>(node_off[n])), nw(static_cast<size_t>(node_off[n]));
  std::vector<int64_t> box_len(n);
  parallel_for(n, [&](size_t b0, size_t b1) {
    for (size_t i = b0; i < b1; ++i) {
      const fh_filament &f = F[i];
      ProjMeta &pm = metas[i];
      const H3 axes[3] = {h3(f.dir), h3(f.wdir), h3(f.hdir)};
      const double extents[3] = {f.length, f.width, f.height};
      const H3 corner = (h3(f.start) - axes[1] * (0.5 * f.width)) - axes[2] * (0.5 * f.height);
      pm.corner[0] = corner.x;
      pm.corner[1] = corner.y;
      pm.corner[2] = corner.z;
      for (int a = 0; a < 3; ++a) {
        pm.axes[3 * a] = axes[a].x;
        pm.axes[3 * a + 1] = axes[a].y;
        pm.axes[3 * a + 2] = axes[a].z;
      }
      pm.inv_area = 1.0 / (f.width * f.height);
      pm.node_off = static_cast<int32_t>(node_off[i]);
      int64_t q = node_off[i];
      for (int d = 0; d < 3; ++d) {
        double pc = std::ceil(extents[d] / h);
        if (pc < 1.0) pc = 1.0;
        const int64_t pieces = static_cast<int64_t>(pc);
        const double piece = extents[d] / static_cast<double>(pieces);
        const int o = pm.count[d] / static_cast<int32_t>(pieces);
        for (int64_t p = 0; p < pieces; ++p)
          for (int k = 0; k < o; ++k) {
            const double a = static_cast<double>(p) * piece, b = static_cast<double>(p + 1) * piece;
            const double mid = 0.5 * (a + b), half = 0.5 * (b - a);
            nx[q] = mid + half * rule_x[o][k];
            nw[q] = half * rule_w[o][k];
            ++q;
          }
      }
      // the box of the stencils
      int64_t blo[3] = {INT64_MAX, INT64_MAX, INT64_MAX}, bhi[3] = {0, 0, 0};
      const double *n0 = &nx[node_off[i]], *n1 = n0 + pm.count[0], *n2 = n1 + pm.count[1];
      for (int a = 0; a < pm.count[0]; ++a)
        for (int b = 0; b < pm.count[1]; ++b)
          for (int c = 0; c < pm.count[2]; ++c) {
            double r[3];
            for (int k = 0; k < 3; ++k)
              r[k] = ((pm.corner[k] + pm.axes[k] * n0[a]) + pm.axes[3 + k] * n1[b]) + pm.axes[6 + k] * n2[c];
            for (int ax = 0; ax < 3; ++ax) {
              double t;
              const int64_t first = stencil_first(g, ax, r[ax], t);
              blo[ax] = std::min(blo[ax], first);
              bhi[ax] = std::max(bhi[ax], first + g.m - 1);
            }
          }
      for (int ax = 0; ax < 3; ++ax) {
        pm.box_lo[ax] = blo[ax];
        pm.box_dims[ax] = static_cast<int32_t>(bhi[ax] - blo[ax] + 1);
      }
      box_len[i] = static_cast<int64_t>(pm.box_dims[0]) * pm.box_dims[1] * pm.box_dims[2];
    }
  });
  int64_t box_total = 0;
  for (size_t i = 0; i < n; ++i) {
    if (box_len[i] > (int64_t{1} << 30))
      return set_error(FH_E_PFFT, "the projection of filament " + std::to_string(i) +
                                      " covers more than 2^30 grid points",
                       static_cast<int64_t>(i), box_len[i], 0.0, 2);
    metas[i].out_off = box_total;
    box_total += box_len[i];
  }
  std::vector<double> box_w(static_cast<size_t>(box_total));
  std::vector<uint8_t> box_t(static_cast<size_t>(box_total));
  {
    DeviceArray<ProjMeta> d_meta;
    DeviceArray<double> d_nx, d_nw, d_bw;
    DeviceArray<uint8_t> d_bt;
    if ((st = d_meta.allocate(n)) || (st = d_nx.allocate(nx.size())) ||
        (st = d_nw.allocate(nw.size())) || (st = d_bw.allocate(static_cast<size_t>(box_total))) ||
        (st = d_bt.allocate(static_cast<size_t>(box_total))))
      return st;
    FH_HIP_TRY(hipMemcpyAsync(d_meta.ptr, metas.data(), n * sizeof(ProjMeta), hipMemcpyHostToDevice, stream));
    FH_HIP_TRY(hipMemcpyAsync(d_nx.ptr, nx.data(), nx.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    FH_HIP_TRY(hipMemcpyAsync(d_nw.ptr, nw.data(), nw.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    project_kernel<<<static_cast<unsigned>(n), 256, 0, stream>>>(d_meta.ptr, d_nx.ptr, d_nw.ptr, g,
                                                                 d_bw.ptr, d_bt.ptr);
    FH_HIP_TRY(hipGetLastError());
    FH_HIP_TRY(hipMemcpyAsync(box_w.data(), d_bw.ptr, box_w.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    FH_HIP_TRY(hipMemcpyAsync(box_t.data(), d_bt.ptr, box_t.size(), hipMemcpyDeviceToHost, stream));
    FH_HIP_TRY(hipStreamSynchronize(stream));
  }
  // compaction: touched points in ascending grid index, entry boxes
  std::vector<int64_t> poff(n + 1, 0);
  for (size_t i = 0; i < n; ++i) {
    int64_t c = 0;
    for (int64_t q = 0; q < box_len[i]; ++q) c += box_t[static_cast<size_t>(metas[i].out_off + q)];
    poff[i + 1] = poff[i] + c;
  }
  const size_t E = static_cast<size_t>(poff[n]);
  op.proj_entries = E;
  std::vector<uint32_t> pidx(E);
  std::vector<double> pw(E);
  std::vector<int64_t> elo(3 * n), ehi(3 * n);
  parallel_for(n, [&](size_t b0, size_t b1) {
    for (size_t i = b0; i < b1; ++i) {
      const ProjMeta &pm = metas[i];
      int64_t e = poff[i];
      int64_t lo_[3] = {INT64_MAX, INT64_MAX, INT64_MAX}, hi_[3] = {0, 0, 0};
      for (int x = 0; x < pm.box_dims[0]; ++x)
        for (int y = 0; y < pm.box_dims[1]; ++y)
          for (int z = 0; z < pm.box_dims[2]; ++z) {
            const int64_t cell = (static_cast<int64_t>(x) * pm.box_dims[1] + y) * pm.box_dims[2] + z;
            if (!box_t[static_cast<size_t>(pm.out_off + cell)]) continue;
            const int64_t p[3] = {pm.box_lo[0] + x, pm.box_lo[1] + y, pm.box_lo[2] + z};
            pidx[static_cast<size_t>(e)] = static_cast<uint32_t>((p[0] * g.dims[1] + p[1]) * g.dims[2] + p[2]);
            pw[static_cast<size_t>(e)] = box_w[static_cast<size_t>(pm.out_off + cell)];
            ++e;
            for (int a = 0; a < 3; ++a) {
              lo_[a] = std::min(lo_[a], p[a]);
              hi_[a] = std::max(hi_[a], p[a]);
            }
          }
      for (int a = 0; a < 3; ++a) {
        elo[3 * i + a] = lo_[a];
        ehi[3 * i + a] = hi_[a];
      }
    }
  });
  std::vector<double>().swap(box_w);
  std::vector<uint8_t>().swap(box_t);
  // transpose (mod.rs:395-417): per grid point, filaments in order
  std::vector<int64_t> soff(grid_total + 1, 0);
  for (size_t e = 0; e < E; ++e) soff[pidx[e] + 1]++;
  for (size_t q = 0; q < grid_total; ++q) soff[q + 1] += soff[q];
  std::vector<uint32_t> sfil(E);
  std::vector<double> sw(E);
  {
    std::vector<int64_t> fill(soff.begin(), soff.end() - 1);
    for (size_t i = 0; i < n; ++i)
      for (int64_t e = poff[i]; e < poff[i + 1]; ++e) {
        const int64_t slot = fill[pidx[static_cast<size_t>(e)]]++;
        sfil[static_cast<size_t>(slot)] = static_cast<uint32_t>(i);
        sw[static_cast<size_t>(slot)] = pw[static_cast<size_t>(e)];
      }
  }
  std::vector<double> dirs(3 * n);
  for (size_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) dirs[3 * i + c] = F[i].dir[c];
  op.ncomp = 0;
  for (int c = 0; c < 3; ++c) {
    bool any = false;
    for (size_t i = 0; i < n && !any; ++i) any = dirs[3 * i + c] != 0.0;
    if (any) op.comps[op.ncomp++] = c;
  }
  if ((st = op.proj_idx.allocate(std::max<size_t>(E, 1))) || (st = op.proj_w.allocate(std::max<size_t>(E, 1))) ||
      (st = op.spread_off.allocate(grid_total + 1)) || (st = op.spread_fil.allocate(std::max<size_t>(E, 1))) ||
      (st = op.spread_w.allocate(std::max<size_t>(E, 1))) || (st = op.dirs.allocate(3 * n)))
    return st;
  FH_HIP_TRY(hipMemcpyAsync(op.proj_off.ptr, poff.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, stream));
  FH_HIP_TRY(hipMemcpyAsync(op.proj_idx.ptr, pidx.data(), E * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
  FH_HIP_TRY(hipMemcpyAsync(op.proj_w.ptr, pw.data(), E * sizeof(double), hipMemcpyHostToDevice, stream));
  FH_HIP_TRY(hipMemcpyAsync(op.spread_off.ptr, soff.data(), (grid_total + 1) * sizeof(int64_t), hipMemcpyHostToDevice, stream));
  FH_HIP_TRY(hipMemcpyAsync(op.spread_fil.ptr, sfil.data(), E * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
  FH_HIP_TRY(hipMemcpyAsync(op.spread_w.ptr, sw.data(), E * sizeof(double), hipMemcpyHostToDevice, stream));
  FH_HIP_TRY(hipMemcpyAsync(op.dirs.ptr, dirs.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice, stream));
  FH_HIP_TRY(hipStreamSynchronize(stream));
  std::vector<uint32_t>().swap(sfil);
  std::vector<double>().swap(sw);
  std::vector<int64_t>().swap(soff);

  // ---- near pairs (near.rs:38-109), host ----
  std::vector<HBar> bars(n);
  std::vector<H3> blo(n), bhi(n);
  for (size_t i = 0; i < n; ++i) {
    bars[i] = bar_of(F[i]);
    H3 reach{0.0, 0.0, 0.0};
    for (int k = 0; k < 3; ++k) reach = reach + habs(bars[i].axes[k]) * bars[i].half[k];
    blo[i] = bars[i].centre - reach;
    bhi[i] = bars[i].centre + reach;
  }
  H3 glo = blo[0], ghi = bhi[0];
  for (size_t i = 0; i < n; ++i) {
    glo = {std::fmin(glo.x, blo[i].x), std::fmin(glo.y, blo[i].y), std::fmin(glo.z, blo[i].z)};
    ghi = {std::fmax(ghi.x, bhi[i].x), std::fmax(ghi.y, bhi[i].y), std::fmax(ghi.z, bhi[i].z)};
  }
  const H3 gext = ghi - glo;
  const double e3[3] = {gext.x, gext.y, gext.z}, l3[3] = {glo.x, glo.y, glo.z};
  double size = std::fmax(std::fmax(radius, std::fmax(std::fmax(gext.x, gext.y), gext.z) * 1e-9),
                          2.2250738585072014e-308);
  size_t bdims[3];
  for (;;) {
    size_t prod = 1;
    for (int a = 0; a < 3; ++a) {
      bdims[a] = static_cast<size_t>(std::floor(e3[a] / size)) + 1;
      prod *= bdims[a];
    }
    if (prod <= 8 * n + 64) break;
    size *= 1.5;
  }
  auto cell = [&](double x, int a) {
    return static_cast<size_t>(std::fmin(std::fmax(std::floor((x - l3[a]) / size), 0.0),
                                         static_cast<double>(bdims[a] - 1)));
  };
  const size_t nbins = bdims[0] * bdims[1] * bdims[2];
  std::vector<size_t> bstart(nbins + 1, 0);
  auto bin_range = [&](H3 a, H3 b, size_t f0[3], size_t f1[3]) {
    f0[0] = cell(a.x, 0), f0[1] = cell(a.y, 1), f0[2] = cell(a.z, 2);
    f1[0] = cell(b.x, 0), f1[1] = cell(b.y, 1), f1[2] = cell(b.z, 2);
  };
  for (size_t i = 0; i < n; ++i) {
    size_t f0[3], f1[3];
    bin_range(blo[i], bhi[i], f0, f1);
    for (size_t a = f0[0]; a <= f1[0]; ++a)
      for (size_t b = f0[1]; b <= f1[1]; ++b)
        for (size_t c = f0[2]; c <= f1[2]; ++c) bstart[(a * bdims[1] + b) * bdims[2] + c + 1]++;
  }
  for (size_t q = 0; q < nbins; ++q) bstart[q + 1] += bstart[q];
  std::vector<uint32_t> items(bstart[nbins]);
  {
    std::vector<size_t> fillp(bstart.begin(), bstart.end() - 1);
    for (size_t i = 0; i < n; ++i) {
      size_t f0[3], f1[3];
      bin_range(blo[i], bhi[i], f0, f1);
      for (size_t a = f0[0]; a <= f1[0]; ++a)
        for (size_t b = f0[1]; b <= f1[1]; ++b)
          for (size_t c = f0[2]; c <= f1[2]; ++c)
            items[fillp[(a * bdims[1] + b) * bdims[2] + c]++] = static_cast<uint32_t>(i);
    }
  }
  // partners: near pairs j >= i, not orthogonal (near.rs:221-233)
  std::vector<std::vector<uint32_t>> partners(n);
  parallel_for(n, [&](size_t b0, size_t b1) {
    std::vector<uint32_t> found;
    for (size_t i = b0; i < b1; ++i) {
      size_t f0[3], f1[3];
      bin_range(blo[i] - H3{radius, radius, radius}, bhi[i] + H3{radius, radius, radius}, f0, f1);
      found.clear();
      for (size_t a = f0[0]; a <= f1[0]; ++a)
        for (size_t b = f0[1]; b <= f1[1]; ++b)
          for (size_t c = f0[2]; c <= f1[2]; ++c) {
            const size_t q = (a * bdims[1] + b) * bdims[2] + c;
            for (size_t e = bstart[q]; e < bstart[q + 1]; ++e)
              if (items[e] >= i) found.push_back(items[e]);
          }
      std::sort(found.begin(), found.end());
      found.erase(std::unique(found.begin(), found.end()), found.end());
      const H3 di = h3(F[i].dir);
      auto &out = partners[i];
      for (uint32_t other : found) {
        if (!(other == i || separation(bars[i], bars[other]) < radius)) continue;
        if (std::fabs(dot(di, h3(F[other].dir))) >= 1e-14) out.push_back(other);
      }
    }
  });
  std::vector<uint32_t>().swap(items);
  // ---- precorrection (near.rs:213-313) ----
  std::vector<int64_t> pair_off(n + 1, 0);
  for (size_t i = 0; i < n; ++i) pair_off[i + 1] = pair_off[i] + static_cast<int64_t>(partners[i].size());
  const size_t P = static_cast<size_t>(pair_off[n]);
  std::vector<uint2> pair_list(P);
  std::vector<int64_t> pj(P);
  for (size_t i = 0; i < n; ++i)
    for (size_t q = 0; q < partners[i].size(); ++q) {
      pair_list[static_cast<size_t>(pair_off[i]) + q] = make_uint2(static_cast<uint32_t>(i), partners[i][q]);
      pj[static_cast<size_t>(pair_off[i]) + q] = partners[i][q];
    }
  // row tasks and the cost rule
  std::vector<RowTask> direct_tasks, box_tasks;
  int64_t scratch_total = 0;
  int box_lds_cells = 0;
  for (size_t i = 0; i < n; ++i) {
    if (partners[i].empty()) continue;
    RowTask t{};
    t.row = static_cast<int64_t>(i);
    t.partner_off = pair_off[i];
    t.partners = static_cast<int32_t>(partners[i].size());
    int64_t plo[3] = {INT64_MAX, INT64_MAX, INT64_MAX}, phi_[3] = {0, 0, 0};
    int64_t partner_entries = 0;
    for (uint32_t j : partners[i]) {
      for (int a = 0; a < 3; ++a) {
        plo[a] = std::min(plo[a], elo[3 * j + a]);
        phi_[a] = std::max(phi_[a], ehi[3 * j + a]);
      }
      partner_entries += poff[j + 1] - poff[j];
    }
    int64_t blen = 1;
    for (int a = 0; a < 3; ++a) {
      t.box_lo[a] = plo[a];
      t.box_dims[a] = static_cast<int32_t>(phi_[a] - plo[a] + 1);
      blen *= t.box_dims[a];
    }
    // near.rs:352-356: the cheaper of the two evaluations
    const int64_t src = poff[i + 1] - poff[i];
    const int64_t direct_cost = src * partner_entries, box_cost = src * blen + partner_entries;
    if (direct_cost <= box_cost) {
      direct_tasks.push_back(t);
      continue;
    }
    t.use_box = 1;
    if (blen > kBoxLds) {
      t.scratch_off = scratch_total;
      scratch_total += blen;
    } else {
      box_lds_cells = std::max(box_lds_cells, static_cast<int>(blen));
    }
    box_tasks.push_back(t);
  }
  // the offsets a row meets: its own entries against its partners' box
  // (near.rs:305-315)
  int64_t reach[3] = {0, 0, 0};
  for (const std::vector<RowTask> *list : {&direct_tasks, &box_tasks})
    for (const RowTask &t : *list) {
      const size_t i = static_cast<size_t>(t.row);
      for (int a = 0; a < 3; ++a) {
        const int64_t plo = t.box_lo[a], phi_ = t.box_lo[a] + t.box_dims[a] - 1;
        const int64_t r1 = phi_ - std::min(elo[3 * i + a], phi_), r2 = ehi[3 * i + a] - std::min(plo, ehi[3 * i + a]);
        reach[a] = std::max(reach[a], std::max(r1, r2));
      }
    }
  const int64_t table_len = (2 * reach[0] + 1) * (2 * reach[1] + 1) * (2 * reach[2] + 1);
  const bool use_table = table_len <= (int64_t{1} << 26);
  std::vector<double> approx(P), exact(P);
  if (P > 0) {
    DeviceArray<fh_filament> d_f;
    DeviceArray<uint2> d_pairs;
    DeviceArray<double> d_exact, d_approx, d_scratch;
    DeviceArray<int64_t> d_pj;
    DeviceArray<RowTask> d_direct, d_box;
    if ((st = d_f.allocate(n)) || (st = d_pairs.allocate(P)) || (st = d_exact.allocate(P)) ||
        (st = d_approx.allocate(P)) || (st = d_pj.allocate(P)) ||
        (st = d_direct.allocate(std::max<size_t>(direct_tasks.size(), 1))) ||
        (st = d_box.allocate(std::max<size_t>(box_tasks.size(), 1))) ||
        (st = d_scratch.allocate(static_cast<size_t>(std::max<int64_t>(scratch_total, 1)))))
      return st;
    FH_HIP_TRY(hipMemcpyAsync(d_f.ptr, F.data(), n * sizeof(fh_filament), hipMemcpyHostToDevice, stream));
    FH_HIP_TRY(hipMemcpyAsync(d_pairs.ptr, pair_list.data(), P * sizeof(uint2), hipMemcpyHostToDevice, stream));
    FH_HIP_TRY(hipMemcpyAsync(d_pj.ptr, pj.data(), P * sizeof(int64_t), hipMemcpyHostToDevice, stream));
    if (!direct_tasks.empty())
      FH_HIP_TRY(hipMemcpyAsync(d_direct.ptr, direct_tasks.data(), direct_tasks.size() * sizeof(RowTask),
                                hipMemcpyHostToDevice, stream));
    if (!box_tasks.empty())
      FH_HIP_TRY(hipMemcpyAsync(d_box.ptr, box_tasks.data(), box_tasks.size() * sizeof(RowTask),
                                hipMemcpyHostToDevice, stream));
    if ((st = fill_pairs(d_f.ptr, n, d_pairs.ptr, P, d_exact.ptr, stream))) return st;
    DeviceArray<double> d_table;
    KernelTable T{nullptr, static_cast<int32_t>(reach[0]), static_cast<int32_t>(reach[1]),
                  static_cast<int32_t>(reach[2]), static_cast<int32_t>(2 * reach[1] + 1),
                  static_cast<int32_t>(2 * reach[2] + 1)};
    if (use_table) {
      if ((st = d_table.allocate(static_cast<size_t>(table_len)))) return st;
      kernel_table_kernel<<<blocks_for(table_len), 256, 0, stream>>>(d_table.ptr, T.r0, T.r1, T.r2, h);
      FH_HIP_TRY(hipGetLastError());
      T.t = d_table.ptr;
    }
    if (!direct_tasks.empty()) {
      if (use_table)
        interactions_direct_kernel<true><<<static_cast<unsigned>(direct_tasks.size()), 256, 0, stream>>>(
            d_direct.ptr, op.proj_off.ptr, op.proj_idx.ptr, op.proj_w.ptr, d_pj.ptr, g, T, d_approx.ptr);
      else
        interactions_direct_kernel<false><<<static_cast<unsigned>(direct_tasks.size()), 256, 0, stream>>>(
            d_direct.ptr, op.proj_off.ptr, op.proj_idx.ptr, op.proj_w.ptr, d_pj.ptr, g, T, d_approx.ptr);
    }
    if (!box_tasks.empty()) {
      const size_t lds = static_cast<size_t>(box_lds_cells) * sizeof(double);
      if (use_table)
        interactions_box_kernel<true><<<static_cast<unsigned>(box_tasks.size()), 512, lds, stream>>>(
            d_box.ptr, op.proj_off.ptr, op.proj_idx.ptr, op.proj_w.ptr, d_pj.ptr, g, T, box_lds_cells,
            d_scratch.ptr, d_approx.ptr);
      else
        interactions_box_kernel<false><<<static_cast<unsigned>(box_tasks.size()), 512, lds, stream>>>(
            d_box.ptr, op.proj_off.ptr, op.proj_idx.ptr, op.proj_w.ptr, d_pj.ptr, g, T, box_lds_cells,
            d_scratch.ptr, d_approx.ptr);
    }
    FH_HIP_TRY(hipGetLastError());
    FH_HIP_TRY(hipMemcpyAsync(exact.data(), d_exact.ptr, P * sizeof(double), hipMemcpyDeviceToHost, stream));
    FH_HIP_TRY(hipMemcpyAsync(approx.data(), d_approx.ptr, P * sizeof(double), hipMemcpyDeviceToHost, stream));
    FH_HIP_TRY(hipStreamSynchronize(stream));
  }
  // C_ij = exact - mu0/4pi cos approx, mirrored into full rows (near.rs:264-300)
  std::vector<int64_t> cnt(n, 0);
  for (size_t i = 0; i < n; ++i)
    for (uint32_t j : partners[i]) {
      cnt[i]++;
      if (j != i) cnt[j]++;
    }
  std::vector<int64_t> noff(n + 1, 0);
  for (size_t i = 0; i < n; ++i) noff[i + 1] = noff[i] + cnt[i];
  const size_t NE = static_cast<size_t>(noff[n]);
  op.near_entries = NE;
  std::vector<uint32_t> ncol(NE);
  std::vector<double> nval(NE);
  std::copy(noff.begin(), noff.end() - 1, cnt.begin());
  for (size_t i = 0; i < n; ++i) {
    const H3 di = h3(F[i].dir);
    for (size_t q = 0; q < partners[i].size(); ++q) {
      const uint32_t j = partners[i][q];
      const size_t k = static_cast<size_t>(pair_off[i]) + q;
      const double c = exact[k] - 1.0e-7 * dot(di, h3(F[j].dir)) * approx[k];
      ncol[static_cast<size_t>(cnt[i])] = j;
      nval[static_cast<size_t>(cnt[i]++)] = c;
      if (j != i) {
        ncol[static_cast<size_t>(cnt[j])] = static_cast<uint32_t>(i);
        nval[static_cast<size_t>(cnt[j]++)] = c;
      }
    }
  }
  if ((st = op.near_col.allocate(std::max<size_t>(NE, 1))) || (st = op.near_val.allocate(std::max<size_t>(NE, 1))))
    return st;
  FH_HIP_TRY(hipMemcpyAsync(op.near_off.ptr, noff.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, stream));
  FH_HIP_TRY(hipMemcpyAsync(op.near_col.ptr, ncol.data(), NE * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
  FH_HIP_TRY(hipMemcpyAsync(op.near_val.ptr, nval.data(), NE * sizeof(double), hipMemcpyHostToDevice, stream));

  // ---- convolver (fft.rs:91-140) ----
  {
    bool fits = false;
    const bool single = prm.fft_precision == 1;
    if ((st = op.conv.init(g.dims, op.padded, h, single, kBatchSources / 2, stream, fits))) return st;
    if (fits) {
      op.use_conv = true;
      op.fft_kind = single ? 1 : 0;
      FH_HIP_TRY(hipStreamSynchronize(stream));
      return FH_OK;
    }
  }
  // a padded axis too long for one LDS line: the whole padded grid through
  // global-memory Stockham passes
  op.fft_kind = 2;
  const int64_t M = op.padded_total();
  if ((st = op.buffer.allocate(static_cast<size_t>(M))) || (st = op.spectrum.allocate(static_cast<size_t>(M))) ||
      (st = op.work.allocate(static_cast<size_t>(M))))
    return st;
  for (int a = 0; a < 3; ++a)
    if ((st = stockham_axis_plan(op.padded[a], op.axes[a], stream))) return st;
  kernel_grid_kernel<<<blocks_for(M), 256, 0, stream>>>(op.buffer.ptr, g.dims[0], g.dims[1], g.dims[2],
                                                        op.padded[0], op.padded[1], op.padded[2], h);
  if ((st = op.fft(true, stream))) return st;
  spectrum_kernel<<<blocks_for(M), 256, 0, stream>>>(op.buffer.ptr, op.spectrum.ptr, M,
                                                     1.0 / static_cast<double>(M));
  FH_HIP_TRY(hipGetLastError());
  FH_HIP_TRY(hipStreamSynchronize(stream));
  return FH_OK;
}

fh_status PfftOperator::apply(const double *d_x, double *d_y, int V, hipStream_t stream) {
  PfftImpl &op = *impl_;
  const int64_t n = static_cast<int64_t>(op.n);
  if (n == 0 || V <= 0) return FH_OK;
  near_kernel<<<static_cast<unsigned>((n + 3) / 4), 256, 0, stream>>>(
      op.near_off.ptr, op.near_col.ptr, op.near_val.ptr, n, V, d_x, d_y);
  FH_HIP_TRY(hipGetLastError());
  if (!op.has_grid || op.ncomp == 0) return FH_OK;
  fh_status st;
  const size_t need = static_cast<size_t>(V) * op.n;
  if (op.far_cap < need) {
    if ((st = op.far.allocate(need))) return st;
    op.far_cap = need;
  }
  if (op.use_conv && op.packed.count < op.n * kBatchSources)
    if ((st = op.packed.allocate(op.n * kBatchSources))) return st;
  zero_kernel<<<blocks_for(static_cast<int64_t>(need)), 256, 0, stream>>>(op.far.ptr, static_cast<int64_t>(need));
  const int S = V * op.ncomp;  // sources (v, c), v outer
  const GridDesc &g = op.g;
  if (op.use_conv) {
    const int64_t G = op.grid_total();
    for (int s0 = 0; s0 < S; s0 += kBatchSources) {
      SourceSet ss{};
      ss.count = std::min(kBatchSources, S - s0);
      for (int q = 0; q < ss.count; ++q) {
        ss.vec[q] = (s0 + q) / op.ncomp;
        ss.comp[q] = op.comps[(s0 + q) % op.ncomp];
      }
      const int C = (ss.count + 1) / 2;
      if (op.conv.single()) {
        float2 *grid = static_cast<float2 *>(op.conv.grid());
        pack_sources_kernel<<<blocks_for(n), 256, 0, stream>>>(op.dirs.ptr, d_x, n, ss, op.packed.ptr);
        spread_compact_kernel<float2><<<xcd_blocks_for(G), 256, 0, stream>>>(
            grid, op.spread_off.ptr, op.spread_fil.ptr, op.spread_w.ptr, op.packed.ptr, G, ss);
        FH_HIP_TRY(hipGetLastError());
        if ((st = op.conv.convolve(C, stream))) return st;
        gather_compact_kernel<float2><<<xcd_blocks_for(64 * n), 256, 0, stream>>>(
            grid, op.proj_off.ptr, op.proj_idx.ptr, op.proj_w.ptr, op.dirs.ptr, n, G, ss, op.far.ptr);
      } else {
        double2 *grid = static_cast<double2 *>(op.conv.grid());
        pack_sources_kernel<<<blocks_for(n), 256, 0, stream>>>(op.dirs.ptr, d_x, n, ss, op.packed.ptr);
        spread_compact_kernel<double2><<<xcd_blocks_for(G), 256, 0, stream>>>(
            grid, op.spread_off.ptr, op.spread_fil.ptr, op.spread_w.ptr, op.packed.ptr, G, ss);
        FH_HIP_TRY(hipGetLastError());
        if ((st = op.conv.convolve(C, stream))) return st;
        gather_compact_kernel<double2><<<xcd_blocks_for(64 * n), 256, 0, stream>>>(
            grid, op.proj_off.ptr, op.proj_idx.ptr, op.proj_w.ptr, op.dirs.ptr, n, G, ss, op.far.ptr);
      }
      FH_HIP_TRY(hipGetLastError());
    }
    finish_kernel<<<blocks_for(static_cast<int64_t>(need)), 256, 0, stream>>>(d_y, op.far.ptr,
                                                                               static_cast<int64_t>(need));
    FH_HIP_TRY(hipGetLastError());
    return FH_OK;
  }
  const int64_t M = op.padded_total();
  for (int s = 0; s < S; s += 2) {
    const int va = s / op.ncomp, ca = op.comps[s % op.ncomp];
    const int vb = s + 1 < S ? (s + 1) / op.ncomp : -1;
    const int cb = s + 1 < S ? op.comps[(s + 1) % op.ncomp] : 0;
    spread_kernel<<<blocks_for(M), 256, 0, stream>>>(
        op.buffer.ptr, op.spread_off.ptr, op.spread_fil.ptr, op.spread_w.ptr, op.dirs.ptr, d_x, n,
        g.dims[0], g.dims[1], g.dims[2], op.padded[1], op.padded[2], M, va, ca, vb, cb);
    if ((st = op.fft(true, stream))) return st;
    multiply_kernel<<<blocks_for(M), 256, 0, stream>>>(op.buffer.ptr, op.spectrum.ptr, M);
    if ((st = op.fft(false, stream))) return st;
    gather_kernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, stream>>>(
        op.buffer.ptr, op.proj_off.ptr, op.proj_idx.ptr, op.proj_w.ptr, op.dirs.ptr, n, g.dims[1],
        g.dims[2], op.padded[1], op.padded[2], op.far.ptr, va, ca, vb, cb);
    FH_HIP_TRY(hipGetLastError());
  }
  finish_kernel<<<blocks_for(static_cast<int64_t>(need)), 256, 0, stream>>>(d_y, op.far.ptr,
                                                                             static_cast<int64_t>(need));
  FH_HIP_TRY(hipGetLastError());
  return FH_OK;
}

fh_pfft_stats PfftOperator::stats() const {
  const PfftImpl &op = *impl_;
  fh_pfft_stats s{};
  s.filaments = op.n;
  s.grid_spacing = op.has_grid ? op.g.h : 0.0;
  for (int a = 0; a < 3; ++a) {
    s.grid_dims[a] = op.has_grid ? static_cast<uint64_t>(op.g.dims[a]) : 0;
    s.fft_dims[a] = op.has_grid ? static_cast<uint64_t>(op.padded[a]) : 0;
  }
  s.projection_entries = op.proj_entries;
  s.near_entries = op.near_entries;
  const uint64_t grid = op.has_grid ? static_cast<uint64_t>(op.grid_total()) : 0;
  const uint64_t padded = op.has_grid ? static_cast<uint64_t>(op.padded_total()) : 0;
  s.memory_bytes = op.proj_entries * 2 * (4 + 8) + (op.n + 1 + grid + 1) * 8 + op.near_entries * 12 +
                   (op.n + 1) * 8 + op.n * 24 + op.far_cap * 8 +
                   (op.use_conv ? op.conv.bytes() : padded * 24);
  s.fft_kind = op.fft_kind;
  if (op.use_conv) op.conv.line_groups(s.line_group);
  return s;
}

fh_status PfftOperator::copy_data(int64_t *proj_off, uint32_t *proj_idx, double *proj_w,
                                  int64_t *near_off, uint32_t *near_col, double *near_val) const {
  const PfftImpl &op = *impl_;
  const size_t n = op.n;
  if (proj_off) FH_HIP_TRY(hipMemcpy(proj_off, op.proj_off.ptr, (n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
  if (proj_idx && op.proj_entries)
    FH_HIP_TRY(hipMemcpy(proj_idx, op.proj_idx.ptr, op.proj_entries * sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (proj_w && op.proj_entries)
    FH_HIP_TRY(hipMemcpy(proj_w, op.proj_w.ptr, op.proj_entries * sizeof(double), hipMemcpyDeviceToHost));
  if (near_off) FH_HIP_TRY(hipMemcpy(near_off, op.near_off.ptr, (n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
  if (near_col && op.near_entries)
    FH_HIP_TRY(hipMemcpy(near_col, op.near_col.ptr, op.near_entries * sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (near_val && op.near_entries)
    FH_HIP_TRY(hipMemcpy(near_val, op.near_val.ptr, op.near_entries * sizeof(double), hipMemcpyDeviceToHost));
  return FH_OK;
}

}  // namespace fh
