// Mesh reduction l_mesh = M L M^T on gfx950 (fasterhenry/src/solve.rs:
// 1172-1200, congruence) and r_mesh = M diag(R) M^T on the host.
//
// The reference forms row i of M.L as product_i[k] = sum_{e in row i} s_e
// L[k][b_e] (sequential, from 0) and then out[i][j] = sum_{e in row j} s_e
// product_i[b_e] for j >= i, mirrored.  Each output here is computed with
// exactly those sums, so l_mesh is bit-identical to the reference's:
//  * short rows (bundle loops have two entries) -- one output per thread,
//    16 x 16 output tiles of the upper triangle per workgroup, results
//    stored row-contiguous in both triangles through an LDS transpose;
//  * long rows (conductor and port loops) -- product_i materialized once
//    as a vector (one thread per branch k, reading L row-contiguously),
//    then one output per thread.
// L is symmetric bit for bit (the fill writes both triangles from one
// evaluation), so L[b][b'] is read as L[b'][b] wherever that is the
// row-contiguous direction.  HBM-bound: every L entry a tile needs is read
// about once per tile row.
#include <hip/hip_runtime.h>

#include <unordered_map>

#include "fh_internal.hpp"

namespace fh {
namespace {

constexpr int kShortRow = 8;
constexpr int kTile = 16;

__device__ __forceinline__ void tile_coords(long long t, int nt, int &ti, int &tj) {
  double disc = (2.0 * nt + 1.0) * (2.0 * nt + 1.0) - 8.0 * static_cast<double>(t);
  int r = static_cast<int>(((2.0 * nt + 1.0) - __builtin_sqrt(disc)) * 0.5);
  if (r < 0) r = 0;
  if (r > nt - 1) r = nt - 1;
  auto start = [nt](long long row) { return row * nt - row * (row - 1) / 2; };
  while (r > 0 && start(r) > t) --r;
  while (r + 1 < nt && start(r + 1) <= t) ++r;
  ti = r;
  tj = r + static_cast<int>(t - start(r));
}

// Outputs (i, j), i <= j, both rows short, over the rows listed in `rows`
// (a contiguous run of short rows starting at row0, count rows).
__global__ void __launch_bounds__(256) short_short_kernel(
    const double *__restrict__ L, int64_t n, const int64_t *__restrict__ row_ptr,
    const int32_t *__restrict__ branch, const double *__restrict__ sign, int row0, int count,
    int nt, double *__restrict__ out, int64_t loops) {
  __shared__ double vals[kTile][kTile + 1];
  int ti, tj;
  tile_coords(static_cast<long long>(blockIdx.x), nt, ti, tj);
  const int r = threadIdx.x / kTile, c = threadIdx.x % kTile;
  const int li = ti * kTile + r, lj = tj * kTile + c;
  double acc = 0.0;
  const bool active = li < count && lj < count && li <= lj;
  if (active) {
    const int i = row0 + li, j = row0 + lj;
    const int64_t ib = row_ptr[i], ie = row_ptr[i + 1];
    for (int64_t e = row_ptr[j]; e < row_ptr[j + 1]; ++e) {
      const int64_t b = branch[e];
      double prod = 0.0;
      for (int64_t f = ib; f < ie; ++f) prod += sign[f] * L[static_cast<int64_t>(branch[f]) * n + b];
      acc += sign[e] * prod;
    }
    out[static_cast<int64_t>(i) * loops + j] = acc;
  }
  vals[r][c] = acc;
  __syncthreads();
  const int ii = ti * kTile + c, jj = tj * kTile + r;
  if (ii < count && jj < count && ii < jj)
    out[static_cast<int64_t>(row0 + jj) * loops + (row0 + ii)] = vals[c][r];
}

// Entries of a long row taken per batch: their loads are issued together,
// then added in row order, so a thread has kBatch (or 2 kBatch) loads in
// flight instead of one (the sums are latency-bound: a few thousand threads
// each walk a row of thousands of entries).
constexpr int kBatch = 16;

// out[i][j] and out[j][i] for a long row j and the short rows i < j listed
// in short_rows: sum over e in row j (in order) of s_e (sum over f in row i
// of s_f L[b_e][b_f]).  Rows i of two entries (bundle loops) keep both
// columns in registers; longer short rows take the general loop.
__global__ void short_long_kernel(const double *__restrict__ L, int64_t n,
                                  const int64_t *__restrict__ row_ptr,
                                  const int32_t *__restrict__ branch,
                                  const double *__restrict__ sign,
                                  const int32_t *__restrict__ short_rows, int n_short, int j,
                                  double *__restrict__ out, int64_t loops) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_short) return;
  const int i = short_rows[t];
  if (i >= j) return;
  const int64_t ib = row_ptr[i], ie = row_ptr[i + 1];
  const int64_t jb = row_ptr[j], je = row_ptr[j + 1];
  double acc = 0.0;
  if (ie - ib == 2) {
    const int64_t c0 = branch[ib], c1 = branch[ib + 1];
    const double s0 = sign[ib], s1 = sign[ib + 1];
    for (int64_t e0 = jb; e0 < je; e0 += kBatch) {
      double v0[kBatch], v1[kBatch], se[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int64_t e = e0 + u < je ? e0 + u : jb;
        const int64_t b = branch[e];
        se[u] = sign[e];
        v0[u] = L[b * n + c0];
        v1[u] = L[b * n + c1];
      }
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        if (e0 + u < je) {
          double prod = 0.0;
          prod += s0 * v0[u];
          prod += s1 * v1[u];
          acc += se[u] * prod;
        }
      }
    }
  } else {
    for (int64_t e = jb; e < je; ++e) {
      const int64_t b = branch[e];
      double prod = 0.0;
      for (int64_t f = ib; f < ie; ++f) prod += sign[f] * L[b * n + branch[f]];
      acc += sign[e] * prod;
    }
  }
  out[static_cast<int64_t>(i) * loops + j] = acc;
  out[static_cast<int64_t>(j) * loops + i] = acc;
}

// product_i[k] = sum_{e in row i} s_e L[b_e][k], k = 0..n-1.
__global__ void product_kernel(const double *__restrict__ L, int64_t n,
                               const int64_t *__restrict__ row_ptr,
                               const int32_t *__restrict__ branch,
                               const double *__restrict__ sign, int i, double *__restrict__ pv) {
  const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (k >= n) return;
  const int64_t ib = row_ptr[i], ie = row_ptr[i + 1];
  double acc = 0.0;
  for (int64_t e0 = ib; e0 < ie; e0 += 2 * kBatch) {
    double v[2 * kBatch], se[2 * kBatch];
#pragma unroll
    for (int u = 0; u < 2 * kBatch; ++u) {
      const int64_t e = e0 + u < ie ? e0 + u : ib;
      se[u] = sign[e];
      v[u] = L[static_cast<int64_t>(branch[e]) * n + k];
    }
#pragma unroll
    for (int u = 0; u < 2 * kBatch; ++u)
      if (e0 + u < ie) acc += se[u] * v[u];
  }
  pv[k] = acc;
}

// out[i][j] = out[j][i] = sum_{e in row j} s_e pv[b_e] for j >= i.
__global__ void long_row_kernel(const double *__restrict__ pv, const int64_t *__restrict__ row_ptr,
                                const int32_t *__restrict__ branch,
                                const double *__restrict__ sign, int i, int64_t loops,
                                double *__restrict__ out) {
  const int64_t j = i + blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (j >= loops) return;
  double acc = 0.0;
  for (int64_t e = row_ptr[j]; e < row_ptr[j + 1]; ++e) acc += sign[e] * pv[branch[e]];
  out[static_cast<int64_t>(i) * loops + j] = acc;
  out[j * loops + i] = acc;
}

}  // namespace

fh_status congruence_inductance(const LoopBasis &basis, size_t n, const double *d_L,
                                double *d_lmesh, hipStream_t stream) {
  const int64_t loops = static_cast<int64_t>(basis.loops);
  if (loops == 0) return FH_OK;
  DeviceArray<int64_t> d_row_ptr;
  DeviceArray<int32_t> d_branch, d_short;
  DeviceArray<double> d_sign, d_pv;
  fh_status st;
  const size_t nnz = basis.branch.size();
  if ((st = d_row_ptr.allocate(basis.row_ptr.size())) || (st = d_branch.allocate(nnz)) ||
      (st = d_sign.allocate(nnz)) || (st = d_pv.allocate(n)))
    return st;
  FH_HIP_TRY(hipMemcpyAsync(d_row_ptr.ptr, basis.row_ptr.data(), basis.row_ptr.size() * sizeof(int64_t),
                            hipMemcpyHostToDevice, stream));
  FH_HIP_TRY(hipMemcpyAsync(d_branch.ptr, basis.branch.data(), nnz * sizeof(int32_t),
                            hipMemcpyHostToDevice, stream));
  FH_HIP_TRY(hipMemcpyAsync(d_sign.ptr, basis.sign.data(), nnz * sizeof(double),
                            hipMemcpyHostToDevice, stream));
  // Runs of short rows, and the long rows.
  std::vector<int32_t> short_rows, long_rows;
  for (int64_t i = 0; i < loops; ++i) {
    const int64_t len = basis.row_ptr[i + 1] - basis.row_ptr[i];
    (len <= kShortRow ? short_rows : long_rows).push_back(static_cast<int32_t>(i));
  }
  if (!short_rows.empty()) {
    if ((st = d_short.allocate(short_rows.size()))) return st;
    FH_HIP_TRY(hipMemcpyAsync(d_short.ptr, short_rows.data(), short_rows.size() * sizeof(int32_t),
                              hipMemcpyHostToDevice, stream));
  }
  // short x short: contiguous runs (usually one run: the bundle loops).
  size_t a = 0;
  while (a < short_rows.size()) {
    size_t b = a + 1;
    while (b < short_rows.size() && short_rows[b] == short_rows[b - 1] + 1) ++b;
    // Pairs across two different runs are handled as a rectangle below.
    const int row0 = short_rows[a], count = static_cast<int>(b - a);
    const int nt = (count + kTile - 1) / kTile;
    const long long tiles = static_cast<long long>(nt) * (nt + 1) / 2;
    short_short_kernel<<<static_cast<unsigned>(tiles), 256, 0, stream>>>(
        d_L, static_cast<int64_t>(n), d_row_ptr.ptr, d_branch.ptr, d_sign.ptr, row0, count, nt,
        d_lmesh, loops);
    FH_HIP_TRY(hipGetLastError());
    a = b;
  }
  // Short rows in different runs (a long row between them): rare; one
  // output per thread through the long-row path of each later run's rows.
  {
    size_t r0 = 0;
    std::vector<std::pair<int32_t, int32_t>> runs;
    while (r0 < short_rows.size()) {
      size_t r1 = r0 + 1;
      while (r1 < short_rows.size() && short_rows[r1] == short_rows[r1 - 1] + 1) ++r1;
      runs.emplace_back(short_rows[r0], static_cast<int32_t>(r1 - r0));
      r0 = r1;
    }
    for (size_t x = 1; x < runs.size(); ++x)
      for (int32_t q = 0; q < runs[x].second; ++q) {
        const int j = runs[x].first + q;  // treat j as the "long" row of short_long
        const int earlier = runs[x].first - short_rows[0];
        (void)earlier;
        // rows of earlier runs only (i < runs[x].first)
        int n_before = 0;
        while (n_before < static_cast<int>(short_rows.size()) && short_rows[n_before] < runs[x].first)
          ++n_before;
        short_long_kernel<<<(n_before + 63) / 64, 64, 0, stream>>>(
            d_L, static_cast<int64_t>(n), d_row_ptr.ptr, d_branch.ptr, d_sign.ptr, d_short.ptr,
            n_before, j, d_lmesh, loops);
        FH_HIP_TRY(hipGetLastError());
      }
  }
  for (int32_t j : long_rows) {
    if (!short_rows.empty()) {
      const int ns = static_cast<int>(short_rows.size());
      short_long_kernel<<<(ns + 63) / 64, 64, 0, stream>>>(
          d_L, static_cast<int64_t>(n), d_row_ptr.ptr, d_branch.ptr, d_sign.ptr, d_short.ptr, ns,
          j, d_lmesh, loops);
      FH_HIP_TRY(hipGetLastError());
    }
    product_kernel<<<static_cast<unsigned>((n + 63) / 64), 64, 0, stream>>>(
        d_L, static_cast<int64_t>(n), d_row_ptr.ptr, d_branch.ptr, d_sign.ptr, j, d_pv.ptr);
    FH_HIP_TRY(hipGetLastError());
    const int64_t rest = loops - j;
    long_row_kernel<<<static_cast<unsigned>((rest + 255) / 256), 256, 0, stream>>>(
        d_pv.ptr, d_row_ptr.ptr, d_branch.ptr, d_sign.ptr, j, loops, d_lmesh);
    FH_HIP_TRY(hipGetLastError());
  }
  FH_HIP_TRY(hipStreamSynchronize(stream));
  return FH_OK;
}

// solve.rs:728-731: product_i[b] = 0 + s r[b] for each b of row i, so
// out[i][j] = sum over row j's entries (in order) of s_e (s_i(b_e) r[b_e]),
// non-zero only where rows share a branch.
void congruence_resistance(const LoopBasis &basis, const std::vector<double> &r,
                           std::vector<int32_t> &rows, std::vector<int32_t> &cols,
                           std::vector<double> &vals) {
  rows.clear();
  cols.clear();
  vals.clear();
  const size_t loops = basis.loops;
  // rows containing each branch, with the sign there
  std::unordered_map<int32_t, std::vector<std::pair<int32_t, double>>> holders;
  for (size_t i = 0; i < loops; ++i)
    for (int64_t e = basis.row_ptr[i]; e < basis.row_ptr[i + 1]; ++e)
      holders[basis.branch[e]].emplace_back(static_cast<int32_t>(i), basis.sign[e]);
  std::vector<double> acc(loops, 0.0);
  std::vector<char> touched(loops, 0);
  std::vector<int32_t> list;
  for (size_t j = 0; j < loops; ++j) {
    list.clear();
    for (int64_t e = basis.row_ptr[j]; e < basis.row_ptr[j + 1]; ++e) {
      const int32_t b = basis.branch[e];
      const double s = basis.sign[e];
      for (const auto &h : holders[b]) {
        if (static_cast<size_t>(h.first) > j) continue;
        if (!touched[h.first]) {
          touched[h.first] = 1;
          acc[h.first] = 0.0;
          list.push_back(h.first);
        }
        acc[h.first] += s * (0.0 + h.second * r[b]);
      }
    }
    for (int32_t i : list) {
      rows.push_back(i);
      cols.push_back(static_cast<int32_t>(j));
      vals.push_back(acc[i]);
      if (static_cast<size_t>(i) != j) {
        rows.push_back(static_cast<int32_t>(j));
        cols.push_back(i);
        vals.push_back(acc[i]);
      }
      touched[i] = 0;
    }
  }
}

}  // namespace fh
