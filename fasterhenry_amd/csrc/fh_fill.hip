// Partial-inductance fill on gfx950.
//
// Replaces partial_inductance_matrix / evaluate_upper
// (fasterhenry/src/inductance/batch.rs:246-253, 355-372) and the per-pair
// dispatch evaluate (inductance/mod.rs:303-369) with its methods:
// point quadrature (neumann.rs:127-266), the 64-term aligned-bar closed form
// (bar.rs:30-106), the aligned-bar quadrature (aligned.rs:78-232) and
// cross-section-sampled line kernels (neumann.rs:280-325, lines.rs:31-170).
//
// Layout and schedule:
//  * fill_tile_kernel: one 16 x 16 filament-pair tile of the upper triangle
//    per workgroup (256 threads, one pair each); the 2 x 16 filaments are
//    staged in LDS.  Orthogonal pairs and point quadrature up to
//    kInlineEvals kernel evaluations are finished in the tile; results go
//    out through an LDS transpose so both the (i, j) and the mirrored (j, i)
//    stores are row-contiguous.  Every other pair is appended to a queue.
//  * fill_deferred_kernel: one wavefront per queued pair.  Point quadrature
//    with large clouds runs lanes over the outer cloud; the closed form runs
//    its 64 corner terms on the 64 lanes; the aligned quadrature walks the
//    reference's panel recursion with an explicit stack in LDS, lanes over
//    the quadrature nodes of each leaf panel; sampled filaments run lanes
//    over the sample pairs.
// Every sum is taken in the reference's order (partial terms computed in
// parallel, then added sequentially), so the values are bit-identical to a
// sequential FP64 evaluation with the same log/atan (fh_math.hpp).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "fh_internal.hpp"
#include "fh_math.hpp"
#include "fh_wave.hpp"

namespace fh {
namespace {

constexpr int kMaxRule = 16;         // largest Gauss rule used on the device
constexpr int kPointMaxOrder = 16;   // neumann.rs:34
constexpr int kSampleMaxOrder = 8;   // neumann.rs:40
constexpr int kCloudCapacity = 256;  // neumann.rs:36
constexpr int kCheapPointPairs = 4096;  // inductance/mod.rs:88
constexpr int kFarOrder = 10;        // aligned.rs:69
constexpr int kCornerOrder = 16;     // aligned.rs:71
constexpr double kSnap = 1e-12;      // aligned.rs:74
constexpr int kTile = 16;
constexpr int kInlineEvals = 1024;   // point-quadrature evaluations done in the tile
// Frames of the aligned quadrature's recursion: the frame at stack index k
// has depth k and frames of depth 200 are leaves (the reference's guard), so
// at most 201 are live.  202 frames of 56 bytes plus the wave's 2 KB of
// lane sums fit twelve one-wave workgroups in a CU's 160 KB of LDS (three
// per SIMD, the register budget of the kernels that use it).
constexpr int kStackCapacity = 202;

enum Method : int8_t { kOrthogonal = 0, kBarClosedForm = 1, kAlignedQuadrature = 2,
                       kPointQuadrature = 3, kSampledFilaments = 4, kTruncated = -1 };
enum Kind : uint32_t { kHeavyPoint = 1u, kAligned = 2u, kSampled = 3u };

__constant__ double c_gx[kMaxRule + 1][kMaxRule];
__constant__ double c_gw[kMaxRule + 1][kMaxRule];
__constant__ double c_point_th[kPointMaxOrder];
__constant__ double c_sample_th[kSampleMaxOrder];

// ---------------------------------------------------------------------------
// Host tables: Gauss-Legendre rules (gauss.rs:25-70) and order thresholds
// (gauss.rs:488-495), computed with the host libm once per device.

void legendre(int n, double x, double &p_out, double &p_prev_out) {
  double p_prev = 1.0, p = x;
  for (int j = 1; j < n; ++j) {
    const double jf = static_cast<double>(j);
    const double p_next = ((2.0 * jf + 1.0) * x * p - jf * p_prev) / (jf + 1.0);
    p_prev = p;
    p = p_next;
  }
  p_out = p;
  p_prev_out = p_prev;
}

void host_rule(int n, double *nodes, double *weights) {
  const double nf = static_cast<double>(n);
  for (int k = 0; k < (n + 1) / 2; ++k) {
    double x = std::cos(3.141592653589793 * (static_cast<double>(k) + 0.75) / (nf + 0.5));
    double derivative = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p, p_prev;
      legendre(n, x, p, p_prev);
      derivative = nf * (x * p - p_prev) / (x * x - 1.0);
      const double step = p / derivative;
      x -= step;
      if (std::fabs(step) <= 4.0 * 2.220446049250313e-16 * std::fmax(std::fabs(x), 1e-3)) break;
    }
    if (n % 2 == 1 && k == n / 2) {
      x = 0.0;
      double p, p_prev;
      legendre(n, x, p, p_prev);
      derivative = nf * (x * p - p_prev) / (x * x - 1.0);
    }
    const double weight = 2.0 / ((1.0 - x * x) * derivative * derivative);
    nodes[n - 1 - k] = x;
    nodes[k] = -x;
    weights[n - 1 - k] = weight;
    weights[k] = weight;
  }
}

fh_status upload_tables() {
  static std::mutex mu;
  static std::vector<int> done;  // per device
  int dev = 0;
  FH_HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  if (static_cast<int>(done.size()) > dev && done[dev]) return FH_OK;
  double gx[kMaxRule + 1][kMaxRule] = {}, gw[kMaxRule + 1][kMaxRule] = {};
  for (int n = 1; n <= kMaxRule; ++n) host_rule(n, gx[n], gw[n]);
  double pt[kPointMaxOrder], st[kSampleMaxOrder];
  const double lp = std::log(1.0 / 1e-8), ls = std::log(1.0 / 1e-6);
  for (int n = 1; n <= kPointMaxOrder; ++n) pt[n - 1] = std::sinh(lp / (2.0 * n));
  for (int n = 1; n <= kSampleMaxOrder; ++n) st[n - 1] = std::sinh(ls / (2.0 * n));
  FH_HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_gx), gx, sizeof(gx)));
  FH_HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_gw), gw, sizeof(gw)));
  FH_HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_point_th), pt, sizeof(pt)));
  FH_HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_sample_th), st, sizeof(st)));
  if (static_cast<int>(done.size()) <= dev) done.resize(dev + 1, 0);
  done[dev] = 1;
  return FH_OK;
}

// ---------------------------------------------------------------------------
// Device vector helpers (nalgebra operation order).

struct D3 {
  double x, y, z;
};
__device__ __forceinline__ D3 d3(double x, double y, double z) { return D3{x, y, z}; }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ D3 operator/(D3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross(D3 a, D3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double norm(D3 a) { return __builtin_sqrt(dot(a, a)); }
__device__ __forceinline__ D3 ld3(const double *p) { return {p[0], p[1], p[2]}; }
__device__ __forceinline__ double signum(double x) {
  if (x != x) return x;
  return __builtin_signbit(x) ? -1.0 : 1.0;
}

__device__ __forceinline__ void rule_on(int n, int i, double a, double b, double &x, double &w) {
  const double mid = 0.5 * (a + b), half = 0.5 * (b - a);
  x = mid + half * c_gx[n][i];
  w = half * c_gw[n][i];
}

// gauss.rs:504-517 order_for; 0 = None.
__device__ __forceinline__ int order_for(const double *th, int max_order, double half,
                                         double distance) {
  if (half <= 0.0) return 1;
  if (distance != distance || distance <= 0.0) return 0;
  const double delta = distance / half;
  for (int i = 0; i < max_order; ++i)
    if (delta >= th[i]) return i + 1;
  return 0;
}

// neumann.rs:57-73
struct Bar {
  D3 centre;
  D3 axes[3];
  double half[3];
};

__device__ __forceinline__ Bar make_bar(const fh_filament &f) {
  Bar b;
  b.centre = (ld3(f.start) + ld3(f.end)) * 0.5;
  b.axes[0] = ld3(f.dir);
  b.axes[1] = ld3(f.wdir);
  b.axes[2] = ld3(f.hdir);
  b.half[0] = 0.5 * f.length;
  b.half[1] = 0.5 * f.width;
  b.half[2] = 0.5 * f.height;
  return b;
}

__device__ __forceinline__ double projected_half_extent(const Bar &b, D3 n) {
  double s = 0.0;
  for (int k = 0; k < 3; ++k) s += b.half[k] * __builtin_fabs(dot(n, b.axes[k]));
  return s;
}

// neumann.rs:97-122
__device__ double separation(const Bar &a, const Bar &b) {
  const D3 delta = b.centre - a.centre;
  const double ra = __builtin_sqrt(a.half[0] * a.half[0] + a.half[1] * a.half[1] + a.half[2] * a.half[2]);
  const double rb = __builtin_sqrt(b.half[0] * b.half[0] + b.half[1] * b.half[1] + b.half[2] * b.half[2]);
  const double distance = norm(delta), reach = ra + rb;
  if (distance > 8.0 * reach) return distance - reach;
  double best = -__builtin_inf();
  auto consider = [&](D3 n) {
    const double gap =
        __builtin_fabs(dot(n, delta)) - projected_half_extent(a, n) - projected_half_extent(b, n);
    best = __builtin_fmax(best, gap);
  };
  for (int k = 0; k < 3; ++k) {
    consider(a.axes[k]);
    consider(b.axes[k]);
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const D3 n = cross(a.axes[i], b.axes[j]);
      const double nn = norm(n);
      if (nn > 1e-9) consider(n / nn);
    }
  return best;
}

// neumann.rs:127-134
__device__ __forceinline__ bool point_orders(const Bar &b, double gap, int o[3]) {
  for (int k = 0; k < 3; ++k) {
    o[k] = order_for(c_point_th, kPointMaxOrder, b.half[k], gap);
    if (!o[k]) return false;
  }
  return o[0] * o[1] * o[2] <= kCloudCapacity;
}

// Total order of Rust's f64::total_cmp.
__device__ __forceinline__ long long total_key(double x) {
  long long a = __double_as_longlong(x);
  return a ^ static_cast<long long>(static_cast<unsigned long long>(a >> 63) >> 1);
}

// inductance/mod.rs:231-254: > 0 when a sorts after b.
__device__ int canonical(const fh_filament &a, const fh_filament &b) {
  const double ka[11] = {a.start[0], a.start[1], a.start[2], a.end[0], a.end[1], a.end[2],
                         a.width,    a.height,   a.wdir[0],  a.wdir[1], a.wdir[2]};
  const double kb[11] = {b.start[0], b.start[1], b.start[2], b.end[0], b.end[1], b.end[2],
                         b.width,    b.height,   b.wdir[0],  b.wdir[1], b.wdir[2]};
#pragma unroll
  for (int i = 0; i < 11; ++i) {
    const long long x = total_key(ka[i]), y = total_key(kb[i]);
    if (x != y) return x < y ? -1 : 1;
  }
  return 0;
}

struct Aligned {
  double length[2], width[2], height[2], offset[3];
};

// inductance/mod.rs:280-300
__device__ bool aligned_of(const Bar &a, const Bar &b, Aligned &o) {
  const double kAligned = 1.0 - 1e-12;
  // values, not a choice between addresses: an address picked at run time
  // would put the whole bar in scratch memory
  const double h1 = b.half[1], h2 = b.half[2];
  const bool same = __builtin_fabs(dot(a.axes[1], b.axes[1])) > kAligned;
  if (!same && !(__builtin_fabs(dot(a.axes[1], b.axes[2])) > kAligned)) return false;
  const double wb = same ? h1 : h2, hb = same ? h2 : h1;
  const D3 delta = b.centre - a.centre;
  o.length[0] = 2.0 * a.half[0];
  o.length[1] = 2.0 * b.half[0];
  o.width[0] = 2.0 * a.half[1];
  o.width[1] = 2.0 * wb;
  o.height[0] = 2.0 * a.half[2];
  o.height[1] = 2.0 * hb;
  o.offset[0] = dot(delta, a.axes[0]);
  o.offset[1] = dot(delta, a.axes[1]);
  o.offset[2] = dot(delta, a.axes[2]);
  return true;
}

// Everything the dispatch of inductance/mod.rs:303-356 decides for a pair.
struct PairPlan {
  Bar a, b;
  double cos, gap, area_a, area_b;
  bool parallel, aligned, point;
  int oa[3], ob[3];
  Aligned ab;
  int orthogonal;
};

__device__ void plan_pair(const fh_filament &fa_in, const fh_filament &fb_in, PairPlan &p) {
  const fh_filament *fa = &fa_in, *fb = &fb_in;
  if (canonical(*fa, *fb) > 0) {
    const fh_filament *t = fa;
    fa = fb;
    fb = t;
  }
  const D3 da = ld3(fa->dir), db = ld3(fb->dir);
  p.cos = dot(da, db);
  p.orthogonal = __builtin_fabs(p.cos) < 1e-14;
  if (p.orthogonal) return;
  p.a = make_bar(*fa);
  p.b = make_bar(*fb);
  p.area_a = fa->width * fa->height;
  p.area_b = fb->width * fb->height;
  p.parallel = norm(cross(da, db)) < 1e-6;
  p.aligned = p.parallel ? aligned_of(p.a, p.b, p.ab) : false;
  p.gap = separation(p.a, p.b);
  bool have = point_orders(p.a, p.gap, p.oa);
  if (have) have = point_orders(p.b, p.gap, p.ob);
  if (have && p.aligned) {
    const int pairs = p.oa[0] * p.oa[1] * p.oa[2] * (p.ob[0] * p.ob[1] * p.ob[2]);
    if (pairs > kCheapPointPairs) have = false;
  }
  p.point = have;
}

// Point of cloud of `b` (neumann.rs:160-175), generated incrementally with
// the same operations as the materialized cloud.
// neumann.rs:237-266 for one outer point: inner sum over b's cloud in four
// interleaved lanes, reduced (l0 + l1) + (l2 + l3).
//
// The four lane sums rotate through r0..r3 so that the next term always
// goes to r0 (a select among four accumulators by a lane-varying index costs
// more than the evaluation's divide); r_j holds lane (k + j) mod 4 after k
// terms, which lane_total undoes.
__device__ __forceinline__ double lane_total(double r0, double r1, double r2, double r3, int k) {
  const int n = k & 3;
  const double in0 = n == 0 ? r0 : n == 1 ? r3 : n == 2 ? r2 : r1;
  const double in1 = n == 0 ? r1 : n == 1 ? r0 : n == 2 ? r3 : r2;
  const double in2 = n == 0 ? r2 : n == 1 ? r1 : n == 2 ? r0 : r3;
  const double in3 = n == 0 ? r3 : n == 1 ? r2 : n == 2 ? r1 : r0;
  return (in0 + in1) + (in2 + in3);
}

__device__ double inner_sum(D3 pa, const Bar &b, const int ob[3]) {
  double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
  int k = 0;
  const double s1 = 2.0 * b.half[1], s2 = 2.0 * b.half[2];
  for (int il = 0; il < ob[0]; ++il) {
    double xl, wl;
    rule_on(ob[0], il, -b.half[0], b.half[0], xl, wl);
    const D3 p1 = b.centre + b.axes[0] * xl;
    for (int iw = 0; iw < ob[1]; ++iw) {
      double xw, ww;
      rule_on(ob[1], iw, -0.5, 0.5, xw, ww);
      const D3 p2 = p1 + b.axes[1] * (s1 * xw);
      const double wlw = wl * ww;
      for (int ih = 0; ih < ob[2]; ++ih) {
        double xh, wh;
        rule_on(ob[2], ih, -0.5, 0.5, xh, wh);
        const D3 p3 = p2 + b.axes[2] * (s2 * xh);
        const double w = wlw * wh;
        const double dx = pa.x - p3.x, dy = pa.y - p3.y, dz = pa.z - p3.z;
        const double t = w / __builtin_sqrt(dx * dx + dy * dy + dz * dz);
        const double s = r0 + t;
        r0 = r1;
        r1 = r2;
        r2 = r3;
        r3 = s;
        ++k;
      }
    }
  }
  return lane_total(r0, r1, r2, r3, k);
}

// The height nodes of b's cloud for a height order O2 known at compile time
// (every pair of the point quadrature has O2 = 2 or 3 on the graded
// cross-sections): the offsets axes[2] * (2 half[2] x_h) and weights w_h
// the innermost loop adds, formed once per pair with the same operations.
template <int O2>
struct HeightNodes {
  D3 v[O2];
  double w[O2];
};

template <int O2>
__device__ __forceinline__ HeightNodes<O2> height_nodes(const Bar &b) {
  HeightNodes<O2> h;
  const double s2 = 2.0 * b.half[2];
#pragma unroll
  for (int ih = 0; ih < O2; ++ih) {
    double xh;
    rule_on(O2, ih, -0.5, 0.5, xh, h.w[ih]);
    h.v[ih] = b.axes[2] * (s2 * xh);
  }
  return h;
}

// inner_sum with the height loop unrolled: term ih of a width step goes to
// lane ih mod 4 of the rotation, which advances by O2 per width step.
template <int O2>
__device__ double inner_sum_h(D3 pa, const Bar &b, const int ob[3], const HeightNodes<O2> &h) {
  double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
  const double s1 = 2.0 * b.half[1];
  for (int il = 0; il < ob[0]; ++il) {
    double xl, wl;
    rule_on(ob[0], il, -b.half[0], b.half[0], xl, wl);
    const D3 p1 = b.centre + b.axes[0] * xl;
    for (int iw = 0; iw < ob[1]; ++iw) {
      double xw, ww;
      rule_on(ob[1], iw, -0.5, 0.5, xw, ww);
      const D3 p2 = p1 + b.axes[1] * (s1 * xw);
      const double wlw = wl * ww;
      double acc[4] = {r0, r1, r2, r3};
#pragma unroll
      for (int ih = 0; ih < O2; ++ih) {
        const D3 p3 = p2 + h.v[ih];
        const double w = wlw * h.w[ih];
        const double dx = pa.x - p3.x, dy = pa.y - p3.y, dz = pa.z - p3.z;
        acc[ih & 3] += w / __builtin_sqrt(dx * dx + dy * dy + dz * dz);
      }
      r0 = acc[O2 & 3];
      r1 = acc[(O2 + 1) & 3];
      r2 = acc[(O2 + 2) & 3];
      r3 = acc[(O2 + 3) & 3];
    }
  }
  return lane_total(r0, r1, r2, r3, ob[0] * ob[1] * O2);
}

// Outer cloud point `index` of bar a (il, iw, ih order) and its weight.
__device__ __forceinline__ void cloud_point(const Bar &a, const int oa[3], int index, D3 &p,
                                            double &w) {
  const int ih = index % oa[2];
  const int iw = (index / oa[2]) % oa[1];
  const int il = index / (oa[2] * oa[1]);
  double xl, wl, xw, ww, xh, wh;
  rule_on(oa[0], il, -a.half[0], a.half[0], xl, wl);
  rule_on(oa[1], iw, -0.5, 0.5, xw, ww);
  rule_on(oa[2], ih, -0.5, 0.5, xh, wh);
  p = ((a.centre + a.axes[0] * xl) + a.axes[1] * (2.0 * a.half[1] * xw)) +
      a.axes[2] * (2.0 * a.half[2] * xh);
  w = wl * ww * wh;
}

// neumann.rs:237-266: the outer sum over a's cloud in order; O2 > 0 is b's
// height order, unrolled (0: any order).
template <int O2>
__device__ double point_sum_h(const Bar &a, const int oa[3], const Bar &b, const int ob[3]) {
  HeightNodes<O2 ? O2 : 1> h;
  if constexpr (O2 > 0) h = height_nodes<O2>(b);
  double total = 0.0;
  const double s1 = 2.0 * a.half[1], s2 = 2.0 * a.half[2];
  for (int il = 0; il < oa[0]; ++il) {
    double xl, wl;
    rule_on(oa[0], il, -a.half[0], a.half[0], xl, wl);
    const D3 p1 = a.centre + a.axes[0] * xl;
    for (int iw = 0; iw < oa[1]; ++iw) {
      double xw, ww;
      rule_on(oa[1], iw, -0.5, 0.5, xw, ww);
      const D3 p2 = p1 + a.axes[1] * (s1 * xw);
      const double wlw = wl * ww;
      for (int ih = 0; ih < oa[2]; ++ih) {
        double xh, wh;
        rule_on(oa[2], ih, -0.5, 0.5, xh, wh);
        const D3 p3 = p2 + a.axes[2] * (s2 * xh);
        double inner;
        if constexpr (O2 > 0)
          inner = inner_sum_h<O2>(p3, b, ob, h);
        else
          inner = inner_sum(p3, b, ob);
        total += (wlw * wh) * inner;
      }
    }
  }
  return total;
}

__device__ double point_sum_serial(const Bar &a, const int oa[3], const Bar &b, const int ob[3]) {
  if (ob[2] == 2) return point_sum_h<2>(a, oa, b, ob);
  if (ob[2] == 3) return point_sum_h<3>(a, oa, b, ob);
  return point_sum_h<0>(a, oa, b, ob);
}

// ---------------------------------------------------------------------------
// bar.rs: closed form.

__device__ __forceinline__ double log_term(double c, double x, double r, double rest_sq) {
  if (c == 0.0 || x == 0.0) return 0.0;
  const double argument = x >= 0.0 ? x + r : rest_sq / (r - x);
  return c * x * pm_log(argument);
}

__device__ __forceinline__ double atan_term(double x, double y, double z, double r) {
  const double product = x * y * z;
  if (product == 0.0) return 0.0;
  return -(product * z * z / 6.0) * pm_atan(x * y / (z * r));
}

__device__ double primitive(double x, double y, double z) {
  const double x2 = x * x, y2 = y * y, z2 = z * z;
  const double x4 = x2 * x2, y4 = y2 * y2, z4 = z2 * z2;
  const double r = __builtin_sqrt(x2 + y2 + z2);
  return log_term(y2 * z2 / 4.0 - y4 / 24.0 - z4 / 24.0, x, r, y2 + z2) +
         log_term(x2 * z2 / 4.0 - x4 / 24.0 - z4 / 24.0, y, r, x2 + z2) +
         log_term(x2 * y2 / 4.0 - x4 / 24.0 - y4 / 24.0, z, r, x2 + y2) +
         (x4 + y4 + z4 - 3.0 * (x2 * y2 + y2 * z2 + z2 * x2)) * r / 60.0 +
         atan_term(x, y, z, r) + atan_term(x, z, y, r) + atan_term(y, z, x, r);
}

// Corner coordinate c of one axis: off + sum, off + diff, off - diff,
// off - sum for c = 0..3 (sum, diff = half the sum, difference of the two
// extents), with sign +, -, -, +.  Chosen by value: a lane-dependent index
// into an array would put the array in scratch memory.
__device__ __forceinline__ double corner(double e1, double e2, double off, int c) {
  const double sum = 0.5 * (e1 + e2), diff = 0.5 * (e1 - e2);
  const double t = (c == 0 || c == 3) ? sum : diff;
  return c < 2 ? off + t : off - t;
}
__device__ __forceinline__ double corner_sign(int c) { return (c == 0 || c == 3) ? 1.0 : -1.0; }

// ---------------------------------------------------------------------------
// lines.rs

__device__ __forceinline__ double parallel_primitive(double q, double rho_sq) {
  q = __builtin_fabs(q);
  const double root = __builtin_sqrt(q * q + rho_sq);
  if (q == 0.0) return -root;
  return q * pm_log(q + root) - root;
}

__device__ __forceinline__ bool lines_parallel(double l1, double l2, double offset, double rho,
                                               double &out) {
  const double rho_sq = rho * rho;
  const double smooth = parallel_primitive(offset + l2, rho_sq) - parallel_primitive(offset, rho_sq) -
                        parallel_primitive(offset + l2 - l1, rho_sq) +
                        parallel_primitive(offset - l1, rho_sq);
  const double overlap =
      __builtin_fmax(__builtin_fmin(l1, offset + l2) - __builtin_fmax(offset, 0.0), 0.0);
  if (overlap == 0.0) {
    out = smooth;
    return true;
  }
  if (rho > 0.0) {
    out = smooth - 2.0 * overlap * pm_log(rho);
    return true;
  }
  return false;
}

__device__ __forceinline__ double skew_log_term(double x, double big_r, double a, double res) {
  if (x == 0.0) return 0.0;
  const double argument = a >= 0.0 ? big_r + a : res / (big_r - a);
  if (argument > 0.0) return x * pm_log(argument);
  return 0.0;
}

__device__ double skew_primitive(double s, double t, double d, double c, double sn) {
  const double sin_sq = sn * sn;
  const double versine = sin_sq / (1.0 + __builtin_fabs(c));
  double planar_sq, a_s, a_t;
  if (c >= 0.0) {
    const double diff = s - t;
    planar_sq = diff * diff + 2.0 * s * t * versine;
    a_s = s * versine - diff;
    a_t = t * versine + diff;
  } else {
    const double sum = s + t;
    planar_sq = sum * sum - 2.0 * s * t * versine;
    a_s = sum - s * versine;
    a_t = sum - t * versine;
  }
  const double big_r = __builtin_sqrt(__builtin_fmax(planar_sq + d * d, 0.0));
  const double logs = skew_log_term(s, big_r, a_s, s * s * sin_sq + d * d) +
                      skew_log_term(t, big_r, a_t, t * t * sin_sq + d * d);
  if (d == 0.0 || big_r == 0.0) return logs;
  return logs - (d / sn) * pm_atan((d * d * c + s * t * sin_sq) / (d * big_r * sn));
}

__device__ bool lines_skew(D3 a0, D3 a1, D3 b0, D3 b1, double sin_min, double &out) {
  const double la = norm(a1 - a0), lb = norm(b1 - b0);
  const D3 u = (a1 - a0) / la, v = (b1 - b0) / lb;
  const double c = dot(u, v);
  const double sn = norm(cross(u, v));
  if (sn != sn || sn < sin_min) return false;
  const D3 w0 = a0 - b0;
  const double du = dot(u, w0), dv = dot(v, w0);
  const double sin_sq = sn * sn;
  const double s0 = (c * dv - du) / sin_sq;
  const double t0 = (dv - c * du) / sin_sq;
  double d = norm((w0 + u * s0) - v * t0);
  if (d <= 1e-14 * (la + lb + norm(w0))) d = 0.0;
  const double s_lo = -s0, s_hi = la - s0, t_lo = -t0, t_hi = lb - t0;
  out = skew_primitive(s_hi, t_hi, d, c, sn) - skew_primitive(s_hi, t_lo, d, c, sn) -
        skew_primitive(s_lo, t_hi, d, c, sn) + skew_primitive(s_lo, t_lo, d, c, sn);
  return true;
}

// ---------------------------------------------------------------------------
// aligned.rs pieces

__device__ __forceinline__ double overlap_weight(double e1, double e2, double off, double u) {
  const double hi = __builtin_fmin(0.5 * e1, off + 0.5 * e2 - u);
  const double lo = __builtin_fmax(-0.5 * e1, off - 0.5 * e2 - u);
  return __builtin_fmax(hi - lo, 0.0);
}

__device__ __forceinline__ double integrand(const Aligned &b, double axial_start, double u,
                                            double v) {
  const double weight = overlap_weight(b.width[0], b.width[1], b.offset[1], u) *
                        overlap_weight(b.height[0], b.height[1], b.offset[2], v);
  if (weight == 0.0) return 0.0;
  double kernel;
  if (!lines_parallel(b.length[0], b.length[1], axial_start, pm_hypot(u, v), kernel)) kernel = 0.0;
  return weight * kernel;
}

// aligned.rs:86-101: sorted, deduplicated break points (<= 5).
__device__ int break_points(double e1, double e2, double off, double *out) {
  const double sum = 0.5 * (e1 + e2), diff = 0.5 * __builtin_fabs(e1 - e2);
  double p[5];
  int n = 4;
  const double raw[4] = {off - sum, off - diff, off + diff, off + sum};
  for (int i = 0; i < 4; ++i) p[i] = __builtin_fabs(raw[i]) <= kSnap * sum ? 0.0 : raw[i];
  if (p[0] < 0.0 && p[3] > 0.0) p[n++] = 0.0;
  // insertion sort by total order
  for (int i = 1; i < n; ++i) {
    const double v = p[i];
    int j = i - 1;
    while (j >= 0 && total_key(p[j]) > total_key(v)) {
      p[j + 1] = p[j];
      --j;
    }
    p[j + 1] = v;
  }
  int m = 0;
  for (int i = 0; i < n; ++i)
    if (m == 0 || !(out[m - 1] == p[i])) out[m++] = p[i];
  return m;
}

// Sequential sum of one value per lane, lane order, added to `acc`; every
// lane returns the same total.  Wave-cooperative: the block is one wave.
__device__ __forceinline__ double wave_ordered_sum(double acc, double term, int count,
                                                   double *lds) {
  __syncthreads();
  lds[threadIdx.x] = term;
  __syncthreads();
  for (int q = 0; q < count; ++q) acc += lds[q];
  return acc;
}

// aligned.rs:125-134, lanes over the 10 x 10 nodes.
__device__ double direct_wave(const Aligned &b, double axial_start, double u0, double u1,
                              double v0, double v1, double *lds) {
  double sum = 0.0;
  for (int base = 0; base < kFarOrder * kFarOrder; base += 64) {
    const int q = base + static_cast<int>(threadIdx.x);
    double term = 0.0;
    if (q < kFarOrder * kFarOrder) {
      double uu, wu, vv, wv;
      rule_on(kFarOrder, q / kFarOrder, u0, u1, uu, wu);
      rule_on(kFarOrder, q % kFarOrder, v0, v1, vv, wv);
      term = wu * wv * integrand(b, axial_start, uu, vv);
    }
    const int count = kFarOrder * kFarOrder - base < 64 ? kFarOrder * kFarOrder - base : 64;
    sum = wave_ordered_sum(sum, term, count, lds);
  }
  return sum;
}

// aligned.rs:138-151, lanes over the 16 x 16 Duffy nodes.
__device__ double corner_wave(const Aligned &b, double axial_start, double a, double bb,
                              double *lds) {
  double sum = 0.0;
  for (int base = 0; base < kCornerOrder * kCornerOrder; base += 64) {
    const int q = base + static_cast<int>(threadIdx.x);
    double tau, w_tau, t, w_t;
    rule_on(kCornerOrder, q / kCornerOrder, 0.0, 1.0, tau, w_tau);
    rule_on(kCornerOrder, q % kCornerOrder, 0.0, 1.0, t, w_t);
    const double s = tau * tau * tau;
    const double radial = w_tau * 3.0 * tau * tau * s;
    const double value =
        integrand(b, axial_start, a * s, bb * s * t) + integrand(b, axial_start, a * s * t, bb * s);
    const double term = radial * w_t * value;
    sum = wave_ordered_sum(sum, term, 64, lds);
  }
  return sum * __builtin_fabs(a * bb);
}

struct Frame {
  double u0, u1, v0, v1, at, acc;
  int16_t depth;
  int8_t stage, axis;
};
static_assert(sizeof(Frame) == 56, "Frame layout");

// LDS of a deferred wavefront: the aligned quadrature's stack, or the inner
// cloud (x, y, z, w) of a heavy point quadrature; one pair uses one of them.
union DeferredScratch {
  Frame stack[kStackCapacity];
  double cloud[4 * kCloudCapacity];
};

__device__ __forceinline__ double gap_of(double r0, double r1) {
  if (r0 > 0.0) return r0;
  if (r1 < 0.0) return -r1;
  return 0.0;
}

// aligned.rs:153-199 panel(), the recursion replayed with an explicit stack
// so that every sum has the reference's association.  The stack is shared by
// the wave (all lanes hold the same values and take the same branches).
__device__ double panel_wave(const Aligned &b, double axial_start, double axial_scale, double u0,
                             double u1, double v0, double v1, Frame *stack, double *lds) {
  int sp = 0;
  stack[0] = Frame{u0, u1, v0, v1, 0.0, 0.0, 0, 0, 0};
  sp = 1;
  double ret = 0.0;
  bool have = false;
  for (;;) {
    if (!have) {
      Frame f = stack[sp - 1];
      const double su = f.u1 - f.u0, sv = f.v1 - f.v0;
      int axis = -1;
      double at = 0.0;
      double leaf = 0.0;
      bool is_leaf = true;
      if (su <= 0.0 || sv <= 0.0) {
        leaf = 0.0;
      } else {
        const double distance = pm_hypot(gap_of(f.u0, f.u1), gap_of(f.v0, f.v1));
        const double size = __builtin_fmax(su, sv);
        if (f.depth >= 200) {
          leaf = direct_wave(b, axial_start, f.u0, f.u1, f.v0, f.v1, lds);
        } else if (distance > 0.0) {
          if (size <= distance) {
            leaf = direct_wave(b, axial_start, f.u0, f.u1, f.v0, f.v1, lds);
          } else if (su >= sv) {
            is_leaf = false;
            axis = 0;
            at = 0.5 * (f.u0 + f.u1);
          } else {
            is_leaf = false;
            axis = 1;
            at = 0.5 * (f.v0 + f.v1);
          }
        } else {
          const double target = __builtin_fmin(__builtin_fmin(su, sv), axial_scale);
          if (su > 2.0 * target) {
            is_leaf = false;
            axis = 0;
            at = f.u0 == 0.0 ? target : -target;
          } else if (sv > 2.0 * target) {
            is_leaf = false;
            axis = 1;
            at = f.v0 == 0.0 ? target : -target;
          } else {
            leaf = corner_wave(b, axial_start, f.u0 == 0.0 ? f.u1 : f.u0,
                               f.v0 == 0.0 ? f.v1 : f.v0, lds);
          }
        }
      }
      if (is_leaf || sp >= kStackCapacity) {
        if (!is_leaf) leaf = direct_wave(b, axial_start, f.u0, f.u1, f.v0, f.v1, lds);
        ret = leaf;
        have = true;
        --sp;
      } else {
        stack[sp - 1].stage = 1;
        stack[sp - 1].axis = axis;
        stack[sp - 1].at = at;
        Frame c = f;
        c.depth = f.depth + 1;
        c.stage = 0;
        if (axis == 0) c.u1 = at;
        else c.v1 = at;
        stack[sp++] = c;
      }
    } else {
      if (sp == 0) return ret;
      Frame &p = stack[sp - 1];
      if (p.stage == 1) {
        p.acc = ret;
        p.stage = 2;
        Frame c = p;
        c.depth = p.depth + 1;
        c.stage = 0;
        if (p.axis == 0) c.u0 = p.at;
        else c.v0 = p.at;
        stack[sp++] = c;
        have = false;
      } else {
        ret = p.acc + ret;
        --sp;
      }
    }
  }
}

// aligned.rs:203-232
__device__ double aligned_quadrature_wave(const Aligned &b, Frame *stack, double *lds) {
  const double axial_start = b.offset[0] + 0.5 * (b.length[0] - b.length[1]);
  const double l1 = b.length[0], l2 = b.length[1];
  const double sep[4] = {axial_start + l2, axial_start, axial_start + l2 - l1, axial_start - l1};
  double largest = 0.0;
  for (int i = 0; i < 4; ++i) largest = __builtin_fmax(largest, __builtin_fabs(sep[i]));
  double axial_scale = __builtin_inf();
  for (int i = 0; i < 4; ++i) {
    const double q = __builtin_fabs(sep[i]);
    if (q > 1e-9 * largest) axial_scale = __builtin_fmin(axial_scale, q);
  }
  double us[5], vs[5];
  const int nu = break_points(b.width[0], b.width[1], b.offset[1], us);
  const int nv = break_points(b.height[0], b.height[1], b.offset[2], vs);
  double sum = 0.0;
  for (int i = 0; i + 1 < nu; ++i)
    for (int j = 0; j + 1 < nv; ++j)
      sum += panel_wave(b, axial_start, axial_scale, us[i], us[i + 1], vs[j], vs[j + 1], stack, lds);
  return sum;
}

// inductance/mod.rs:265-276 with bar.rs:84-106 on the lanes.
__device__ double aligned_dispatch_wave(const Aligned &b, int8_t &method, Frame *stack,
                                        double *lds) {
  const int lane = static_cast<int>(threadIdx.x);
  const int i = lane >> 4, j = (lane >> 2) & 3, k = lane & 3;
  const double qx = corner(b.length[0], b.length[1], b.offset[0], i);
  const double qy = corner(b.width[0], b.width[1], b.offset[1], j);
  const double qz = corner(b.height[0], b.height[1], b.offset[2], k);
  const double term = corner_sign(i) * corner_sign(j) * corner_sign(k) * primitive(qx, qy, qz);
  __syncthreads();
  lds[lane] = term;
  __syncthreads();
  double value = 0.0, magnitude = 0.0;
  for (int q = 0; q < 64; ++q) {
    value += lds[q];
    magnitude += __builtin_fabs(lds[q]);
  }
  const double err = value > 0.0 ? 8.0 * 2.220446049250313e-16 * magnitude / value : __builtin_inf();
  if (err <= 1e-10) {
    method = kBarClosedForm;
    return value;
  }
  method = kAlignedQuadrature;
  return aligned_quadrature_wave(b, stack, lds);
}

// ---------------------------------------------------------------------------
// Error record: the lowest failing (i, j) with its KernelError sub-code.

__device__ __forceinline__ void record_error(unsigned long long *err, long long i, long long j,
                                             long long n, int code) {
  const unsigned long long key = static_cast<unsigned long long>((i * n + j) * 4 + code);
  atomicMin(err, key);
}

// ---------------------------------------------------------------------------
// Kernels

__device__ __forceinline__ void tile_coords(long long t, int nt, int &ti, int &tj) {
  // Upper-triangular tile t -> (ti, tj), ti <= tj, rows of length nt - ti.
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

// PairMask::keeps (inductance/batch.rs:76-85): pairs of filaments whose
// groups are not coupled are never evaluated.  A filament without a group
// entry, or a table index out of range, keeps the pair.
struct MaskArgs {
  const int32_t *group;  // null: every pair kept
  int n, groups;
  const uint8_t *table;  // groups x groups
};

__device__ __forceinline__ bool mask_keeps(const MaskArgs &m, int i, int j) {
  if (!m.group || i >= m.n || j >= m.n) return true;
  const int a = m.group[i], b = m.group[j];
  if (a < 0 || b < 0) return true;
  const long long flat = static_cast<long long>(a) * m.groups + b;
  if (flat >= static_cast<long long>(m.groups) * m.groups) return true;
  return m.table[flat] != 0;
}

// Four waves per SIMD (128 registers, a few spilled): the FP64 division and
// square-root chains of the inline quadratures need the other waves to
// cover them (12.0 -> 6.0 ms per fill of the batch problems against two).
__global__ void __launch_bounds__(256, 4) fill_tile_kernel(const fh_filament *__restrict__ F, int n,
                                                        int nt, double *__restrict__ L,
                                                        int8_t *__restrict__ methods,
                                                        uint2 *__restrict__ queue,
                                                        unsigned int *__restrict__ queue_count,
                                                        unsigned long long *__restrict__ err,
                                                        MaskArgs mask) {
  __shared__ fh_filament rows[kTile], cols[kTile];
  __shared__ double vals[kTile][kTile + 1];
  __shared__ int8_t meth[kTile][kTile + 1];
  __shared__ int8_t done[kTile][kTile + 1];
  __shared__ unsigned long long order[kTile * kTile];
  __shared__ int inline_count;
  int ti, tj;
  tile_coords(static_cast<long long>(blockIdx.x), nt, ti, tj);
  const int r = threadIdx.x / kTile, c = threadIdx.x % kTile;
  // Stage the two filament blocks (19 doubles each) in LDS.
  {
    const int per = sizeof(fh_filament) / sizeof(double);
    double *dr = reinterpret_cast<double *>(rows), *dc = reinterpret_cast<double *>(cols);
    for (int q = threadIdx.x; q < kTile * per; q += blockDim.x) {
      const int f = q / per, w = q % per;
      const int gi = ti * kTile + f, gj = tj * kTile + f;
      dr[q] = gi < n ? reinterpret_cast<const double *>(F + gi)[w] : 0.0;
      dc[q] = gj < n ? reinterpret_cast<const double *>(F + gj)[w] : 0.0;
    }
  }
  __syncthreads();
  const int i = ti * kTile + r, j = tj * kTile + c;
  double value = 0.0;
  int8_t method = kOrthogonal;
  bool finished = false;
  bool queued = false;
  uint32_t kind = 0;
  // Sort key of an inline point-quadrature pair: its six orders, so that
  // pairs with the same cloud shapes share a wavefront below; ~0 otherwise.
  unsigned long long key = ~0ull;
  if (i < n && j < n && i <= j && !mask_keeps(mask, i, j)) {
    finished = true;  // truncated: exactly 0, never evaluated
    method = kTruncated;
  } else if (i < n && j < n && i <= j) {
    PairPlan p;
    plan_pair(rows[r], cols[c], p);
    if (p.orthogonal) {
      finished = true;
    } else if (p.point && p.oa[0] * p.oa[1] * p.oa[2] * (p.ob[0] * p.ob[1] * p.ob[2]) <= kInlineEvals) {
      method = kPointQuadrature;
      finished = true;
      key = 0;
      for (int q = 0; q < 3; ++q) key = (key << 5) | static_cast<unsigned>(p.ob[2 - q]);
      for (int q = 0; q < 3; ++q) key = (key << 5) | static_cast<unsigned>(p.oa[2 - q]);
    } else {
      kind = p.point ? kHeavyPoint : (p.aligned ? kAligned : kSampled);
      queued = true;
    }
  }
  {
    const unsigned slot = wave_append(queued, queue_count);
    if (queued) queue[slot] = make_uint2(static_cast<uint32_t>(i) | (kind << 30), static_cast<uint32_t>(j));
  }
  vals[r][c] = value;
  meth[r][c] = method;
  done[r][c] = finished ? 1 : 0;
  // The inline point quadratures, sorted by their orders (bitonic, one
  // element per thread) and dealt out in four contiguous quarters, one per
  // wavefront: lanes of a wavefront then run the same loops instead of the
  // longest of 64 different ones, and the wavefronts get equal shares.
  const int t = static_cast<int>(threadIdx.x);
  order[t] = (key << 8) | static_cast<unsigned>(t);
  if (t == 0) inline_count = 0;
  __syncthreads();
  if (key != ~0ull) atomicAdd(&inline_count, 1);
  for (int k = 2; k <= kTile * kTile; k <<= 1) {
    for (int h = k >> 1; h > 0; h >>= 1) {
      const int partner = t ^ h;
      if (partner > t) {
        const unsigned long long a = order[t], b = order[partner];
        if ((a > b) == ((t & k) == 0)) {
          order[t] = b;
          order[partner] = a;
        }
      }
      __syncthreads();
    }
  }
  {
    const int per = (inline_count + 3) / 4, lane = t % 64, pos = (t / 64) * per + lane;
    const unsigned long long e = lane < per && pos < inline_count ? order[pos] : ~0ull;
    if ((e >> 8) != (~0ull >> 8)) {
      const int slot = static_cast<int>(e & 0xff), rr = slot / kTile, cc = slot % kTile;
      PairPlan p;
      plan_pair(rows[rr], cols[cc], p);
      const double v = 1.0e-7 * (point_sum_serial(p.a, p.oa, p.b, p.ob) * p.cos);
      if (!__builtin_isfinite(v)) record_error(err, ti * kTile + rr, tj * kTile + cc, n, 2);
      vals[rr][cc] = v;
    }
  }
  __syncthreads();
  value = vals[r][c];
  if (finished) {
    L[static_cast<size_t>(i) * n + j] = value;
    if (methods) methods[static_cast<size_t>(i) * n + j] = method;
  }
  // Mirrored store: thread (r, c) writes entry (tj*16 + r, ti*16 + c) = pair
  // (i', j') = (ti*16 + c, tj*16 + r), held at vals[c][r].
  const int ii = ti * kTile + c, jj = tj * kTile + r;
  if (ii < n && jj < n && ii < jj && done[c][r]) {
    L[static_cast<size_t>(jj) * n + ii] = vals[c][r];
    if (methods) methods[static_cast<size_t>(jj) * n + ii] = meth[c][r];
  }
}

// One queued pair on one wavefront (the deferred methods): value
// 1e-7 * integral, its Method, and false for the divergent sampled case.
__device__ void deferred_eval(const fh_filament &fa, const fh_filament &fb, uint32_t kind,
                              double *lds, DeferredScratch &sh, int lane, double &v,
                              int8_t &method, bool &ok) {
  PairPlan p;
  plan_pair(fa, fb, p);
  double integral = 0.0;
  method = kPointQuadrature;
  ok = true;
  if (kind == kHeavyPoint) {
    // neumann.rs:237-266.  b's cloud is formed once into LDS; then four
    // lanes per outer point i each run one of the four interleaved lane
    // sums over it (lane s adds the terms j = s, s + 4, ... in order), and
    // (l0 + l1) + (l2 + l3) is formed across them; the outer sum of
    // w_i * inner_i runs in order of i.
    const int na = p.oa[0] * p.oa[1] * p.oa[2], nb = p.ob[0] * p.ob[1] * p.ob[2];
    double *cx = sh.cloud, *cy = cx + kCloudCapacity, *cz = cy + kCloudCapacity,
           *cw = cz + kCloudCapacity;
    __syncthreads();
    for (int q = lane; q < nb; q += 64) {
      D3 pb;
      double wb;
      cloud_point(p.b, p.ob, q, pb, wb);
      cx[q] = pb.x;
      cy[q] = pb.y;
      cz[q] = pb.z;
      cw[q] = wb;
    }
    __syncthreads();
    for (int base = 0; base < 4 * na; base += 64) {
      const int q = base + lane, i = q >> 2, slot = q & 3;
      double acc = 0.0, wa = 0.0;
      if (i < na) {
        D3 pa;
        cloud_point(p.a, p.oa, i, pa, wa);
        for (int jj = slot; jj < nb; jj += 4) {
          const double dx = pa.x - cx[jj], dy = pa.y - cy[jj], dz = pa.z - cz[jj];
          acc += cw[jj] / __builtin_sqrt(dx * dx + dy * dy + dz * dz);
        }
      }
      // (l0 + l1) + (l2 + l3): each addition is commutative bit for bit
      const double pair = acc + __shfl_xor(acc, 1);
      const double inner = pair + __shfl_xor(pair, 2);
      if (i < na && slot == 0) lds[i] = wa * inner;
    }
    __syncthreads();
    double total = 0.0;
    for (int idx = 0; idx < na; ++idx) total += lds[idx];
    integral = total * p.cos;
  } else if (kind == kAligned) {
    const double in = aligned_dispatch_wave(p.ab, method, sh.stack, lds);
    const double scale = signum(p.cos) / (p.area_a * p.area_b);
    integral = in * scale;
  } else {
    // neumann.rs:280-325, lanes over the sample pairs.
    method = kSampledFilaments;
    int oa[2], ob[2];
    const double halves[4] = {p.a.half[1], p.a.half[2], p.b.half[1], p.b.half[2]};
    int *targets[4] = {&oa[0], &oa[1], &ob[0], &ob[1]};
    for (int t = 0; t < 4; ++t) {
      const int o = order_for(c_sample_th, kSampleMaxOrder, halves[t], p.gap);
      *targets[t] = o ? o : kSampleMaxOrder + (t >= 2 ? 1 : 0);
    }
    const int na = oa[0] * oa[1], nb = ob[0] * ob[1], total_pairs = na * nb;
    const double la = 2.0 * p.a.half[0], lb = 2.0 * p.b.half[0];
    double value = 0.0;
    int bad = 0;
    for (int base = 0; base < total_pairs; base += 64) {
      const int idx = base + lane;
      double term = 0.0;
      if (idx < total_pairs) {
        const int ia = idx / nb, ib = idx % nb;
        double xw, ww, xh, wh;
        rule_on(oa[0], ia / oa[1], -0.5, 0.5, xw, ww);
        rule_on(oa[1], ia % oa[1], -0.5, 0.5, xh, wh);
        const D3 pa = p.a.centre + (p.a.axes[1] * (2.0 * p.a.half[1] * xw) +
                                    p.a.axes[2] * (2.0 * p.a.half[2] * xh));
        const double wa = ww * wh;
        rule_on(ob[0], ib / ob[1], -0.5, 0.5, xw, ww);
        rule_on(ob[1], ib % ob[1], -0.5, 0.5, xh, wh);
        const D3 pb = p.b.centre + (p.b.axes[1] * (2.0 * p.b.half[1] * xw) +
                                    p.b.axes[2] * (2.0 * p.b.half[2] * xh));
        const double wb = ww * wh;
        double kernel = 0.0;
        bool good;
        if (p.parallel) {
          const D3 axis = p.a.axes[0];
          const D3 delta = pb - pa;
          const double along = dot(delta, axis);
          const double rho = norm(delta - axis * along);
          good = lines_parallel(la, lb, along + 0.5 * (la - lb), rho, kernel);
        } else {
          const D3 ha = p.a.axes[0] * p.a.half[0], hb = p.b.axes[0] * p.b.half[0];
          good = lines_skew(pa - ha, pa + ha, pb - hb, pb + hb, 0.0, kernel);
        }
        if (!good) bad = 1;
        term = wa * wb * kernel;
      }
      const int cnt = total_pairs - base < 64 ? total_pairs - base : 64;
      value = wave_ordered_sum(value, term, cnt, lds);
    }
    bad = __syncthreads_or(bad);
    if (bad) ok = false;
    const double direction = p.parallel ? signum(p.cos) : p.cos;
    integral = value * direction;
  }
  v = 1.0e-7 * integral;
}

__global__ void __launch_bounds__(64) fill_deferred_kernel(const fh_filament *__restrict__ F, int n,
                                                           double *__restrict__ L,
                                                           int8_t *__restrict__ methods,
                                                           const uint2 *__restrict__ queue,
                                                           const unsigned int *__restrict__ queue_count,
                                                           unsigned long long *__restrict__ err,
                                                           int skip_aligned) {
  __shared__ double lds[kCloudCapacity];
  __shared__ DeferredScratch scratch_lds;
  const unsigned int count = *queue_count;
  const int lane = static_cast<int>(threadIdx.x);
  for (unsigned int e = blockIdx.x; e < count; e += gridDim.x) {
    const uint2 q = queue[e];
    const uint32_t kind = q.x >> 30;
    if (skip_aligned && kind == kAligned) continue;
    const int i = static_cast<int>(q.x & 0x3fffffffu), j = static_cast<int>(q.y);
    double v;
    int8_t method;
    bool ok;
    deferred_eval(F[i], F[j], kind, lds, scratch_lds, lane, v, method, ok);
    if (lane == 0) {
      if (!ok) {
        record_error(err, i, j, n, 1);
      } else {
        if (!__builtin_isfinite(v)) record_error(err, i, j, n, 2);
        L[static_cast<size_t>(i) * n + j] = v;
        L[static_cast<size_t>(j) * n + i] = v;
        if (methods) {
          methods[static_cast<size_t>(i) * n + j] = method;
          methods[static_cast<size_t>(j) * n + i] = method;
        }
      }
    }
    __syncthreads();
  }
}

// Pair list: out[k] = partial inductance of filaments pairs[k] = (i, j).
// Cheap pairs are finished by one thread each; the rest are queued (pair
// index | kind << 30) for one wavefront each, as in the tile fill.
__global__ void __launch_bounds__(256, 4) pairs_kernel(const fh_filament *__restrict__ F, int n,
                                                    const uint2 *__restrict__ pairs, size_t count,
                                                    double *__restrict__ out,
                                                    uint2 *__restrict__ queue,
                                                    unsigned int *__restrict__ queue_count,
                                                    unsigned long long *__restrict__ err) {
  const size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  bool queued = false;
  uint32_t kind = 0;
  if (k < count) {
    const int i = static_cast<int>(pairs[k].x), j = static_cast<int>(pairs[k].y);
    PairPlan p;
    plan_pair(F[i], F[j], p);
    if (p.orthogonal) {
      out[k] = 0.0;
    } else if (p.point && p.oa[0] * p.oa[1] * p.oa[2] * (p.ob[0] * p.ob[1] * p.ob[2]) <= kInlineEvals) {
      const double value = 1.0e-7 * (point_sum_serial(p.a, p.oa, p.b, p.ob) * p.cos);
      out[k] = value;
      if (!__builtin_isfinite(value)) record_error(err, i, j, n, 2);
    } else {
      kind = p.point ? kHeavyPoint : (p.aligned ? kAligned : kSampled);
      queued = true;
    }
  }
  const unsigned slot = wave_append(queued, queue_count);
  if (queued) queue[slot] = make_uint2(static_cast<uint32_t>(k) | (kind << 30), 0u);
}

__global__ void __launch_bounds__(64) pairs_deferred_kernel(
    const fh_filament *__restrict__ F, int n, const uint2 *__restrict__ pairs,
    double *__restrict__ out, const uint2 *__restrict__ queue,
    const unsigned int *__restrict__ queue_count, unsigned long long *__restrict__ err,
    int skip_aligned) {
  __shared__ double lds[kCloudCapacity];
  __shared__ DeferredScratch scratch_lds;
  const unsigned int count = *queue_count;
  const int lane = static_cast<int>(threadIdx.x);
  for (unsigned int e = blockIdx.x; e < count; e += gridDim.x) {
    const uint2 q = queue[e];
    const uint32_t kind = q.x >> 30;
    if (skip_aligned && kind == kAligned) continue;
    const size_t k = q.x & 0x3fffffffu;
    const int i = static_cast<int>(pairs[k].x), j = static_cast<int>(pairs[k].y);
    double v;
    int8_t method;
    bool ok;
    deferred_eval(F[i], F[j], kind, lds, scratch_lds, lane, v, method, ok);
    if (lane == 0) {
      if (!ok) {
        record_error(err, i, j, n, 1);
      } else {
        if (!__builtin_isfinite(v)) record_error(err, i, j, n, 2);
        out[k] = v;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// The aligned-bar memo (inductance/memo.rs:106-131) as a deduplication of the
// deferred queue: an aligned pair's integral depends only on its nine
// AlignedBars numbers, so pairs with the same bits share one evaluation.
// Keys are hashed, sorted, split into runs of bit-equal keys; one wavefront
// evaluates each run head; every member takes the head's integral and applies
// its own scale.  A shared value is the very bits its own evaluation would
// give.

// Queue entry -> (i, j): tile queues hold (i | kind << 30, j), pair-list
// queues (k | kind << 30, 0) with (i, j) = pairs[k].
__device__ inline void entry_pair(uint2 q, const uint2 *pairs, int &i, int &j, uint32_t &kind, uint32_t &k) {
  kind = q.x >> 30;
  k = q.x & 0x3fffffffu;
  if (pairs) {
    const uint2 pr = pairs[k];
    i = static_cast<int>(pr.x);
    j = static_cast<int>(pr.y);
  } else {
    i = static_cast<int>(k);
    j = static_cast<int>(q.y);
  }
}

// The AlignedBars of a pair as plan_pair forms them (canonical order).
__device__ inline void aligned_bars_of(const fh_filament &fa_in, const fh_filament &fb_in, Aligned &ab) {
  const fh_filament *fa = &fa_in, *fb = &fb_in;
  if (canonical(*fa, *fb) > 0) {
    const fh_filament *t = fa;
    fa = fb;
    fb = t;
  }
  const Bar a = make_bar(*fa), b = make_bar(*fb);
  (void)aligned_of(a, b, ab);
}

constexpr unsigned long long kNoKey = ~0ull;

__global__ void memo_key_kernel(const fh_filament *__restrict__ F, const uint2 *__restrict__ queue,
                                const uint2 *__restrict__ pairs, unsigned count, double *__restrict__ keys,
                                unsigned long long *__restrict__ hash, unsigned *__restrict__ idx) {
  const unsigned e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= count) return;
  idx[e] = e;
  int i, j;
  uint32_t kind, k;
  entry_pair(queue[e], pairs, i, j, kind, k);
  unsigned long long h = kNoKey;
  if (kind == kAligned) {
    Aligned ab;
    aligned_bars_of(F[i], F[j], ab);
    const double w[9] = {ab.length[0], ab.length[1], ab.width[0], ab.width[1], ab.height[0],
                         ab.height[1], ab.offset[0], ab.offset[1], ab.offset[2]};
    h = 0x9E3779B97F4A7C15ull;
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      keys[static_cast<size_t>(e) * 9 + t] = w[t];
      h ^= static_cast<unsigned long long>(__double_as_longlong(w[t]));
      h ^= h >> 30;
      h *= 0xbf58476d1ce4e5b9ull;
      h ^= h >> 27;
      h *= 0x94d049bb133111ebull;
      h ^= h >> 31;
    }
    if (h == kNoKey) h = kNoKey - 1;
  }
  hash[e] = h;
}

__global__ void memo_head_kernel(const unsigned long long *__restrict__ sh, const unsigned *__restrict__ si,
                                 const double *__restrict__ keys, unsigned count, unsigned *__restrict__ head) {
  const unsigned q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= count) return;
  unsigned hd = 0;
  if (sh[q] != kNoKey) {
    hd = 1;
    if (q > 0 && sh[q - 1] == sh[q]) {
      bool same = true;
      for (int t = 0; t < 9; ++t)
        same = same && __double_as_longlong(keys[static_cast<size_t>(si[q]) * 9 + t]) ==
                           __double_as_longlong(keys[static_cast<size_t>(si[q - 1]) * 9 + t]);
      if (same) hd = 0;
    }
  }
  head[q] = hd;
}

__global__ void memo_compact_kernel(const unsigned *__restrict__ head, const unsigned *__restrict__ run,
                                    const unsigned *__restrict__ si, unsigned count, unsigned *__restrict__ uq,
                                    unsigned *__restrict__ ucount) {
  const unsigned q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= count) return;
  if (head[q]) uq[run[q] - 1] = si[q];
  if (q == count - 1) *ucount = run[q];
}

// three waves per SIMD: 3.2 -> 2.8 ms per batch problem against two (the
// deferred kernel, with its larger frame, is faster at two)
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) memo_eval_kernel(const fh_filament *__restrict__ F,
                                                       const uint2 *__restrict__ queue,
                                                       const uint2 *__restrict__ pairs,
                                                       const unsigned *__restrict__ uq,
                                                       const unsigned *__restrict__ ucount,
                                                       double *__restrict__ uin, int8_t *__restrict__ umethod) {
  __shared__ double lds[kCloudCapacity];
  __shared__ Frame stack[kStackCapacity];
  const unsigned count = *ucount;
  for (unsigned u = blockIdx.x; u < count; u += gridDim.x) {
    int i, j;
    uint32_t kind, k;
    entry_pair(queue[uq[u]], pairs, i, j, kind, k);
    Aligned ab;
    aligned_bars_of(F[i], F[j], ab);
    int8_t method;
    const double in = aligned_dispatch_wave(ab, method, stack, lds);
    if (threadIdx.x == 0) {
      uin[u] = in;
      umethod[u] = method;
    }
    __syncthreads();
  }
}

// Every aligned entry: v = 1e-7 (integral * sign(cos) / (area_a area_b)), as
// deferred_eval scales it, into L (both triangles) or the pair list.
__global__ void memo_scatter_kernel(const fh_filament *__restrict__ F, int n, const uint2 *__restrict__ queue,
                                    const uint2 *__restrict__ pairs, const unsigned long long *__restrict__ sh,
                                    const unsigned *__restrict__ si, const unsigned *__restrict__ run,
                                    unsigned count, const double *__restrict__ uin,
                                    const int8_t *__restrict__ umethod, double *__restrict__ L,
                                    int8_t *__restrict__ methods, double *__restrict__ out,
                                    unsigned long long *__restrict__ err) {
  const unsigned q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= count || sh[q] == kNoKey) return;
  int i, j;
  uint32_t kind, k;
  entry_pair(queue[si[q]], pairs, i, j, kind, k);
  const fh_filament &fa = F[i], &fb = F[j];
  const double cos_ = dot(ld3(fa.dir), ld3(fb.dir));
  const double scale = signum(cos_) / ((fa.width * fa.height) * (fb.width * fb.height));
  const double v = 1.0e-7 * (uin[run[q] - 1] * scale);
  if (!__builtin_isfinite(v)) record_error(err, i, j, n, 2);
  if (pairs) {
    out[k] = v;
  } else {
    L[static_cast<size_t>(i) * n + j] = v;
    L[static_cast<size_t>(j) * n + i] = v;
    if (methods) {
      methods[static_cast<size_t>(i) * n + j] = umethod[run[q] - 1];
      methods[static_cast<size_t>(j) * n + i] = umethod[run[q] - 1];
    }
  }
}

__global__ void device_math_kernel(const double *x, size_t n, double *lg, double *at) {
  const size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (k < n) {
    lg[k] = pm_log(x[k]);
    at[k] = pm_atan(x[k]);
  }
}

}  // namespace

fh_status fill_tables_ready() { return upload_tables(); }

namespace {

constexpr unsigned kMemoMin = 256;  // shorter queues are evaluated pair by pair

// Runs the memo over the aligned entries of a deferred queue (count read from
// the device); `used` = false when the queue is too short to bother, and the
// plain deferred kernel must then evaluate the aligned entries too.
fh_status memo_aligned(const fh_filament *d_F, int n, const uint2 *queue, const uint2 *pairs,
                       const unsigned int *d_count, double *d_L, int8_t *d_methods, double *d_out,
                       unsigned long long *err, hipStream_t stream, bool &used) {
  used = false;
  unsigned count = 0;
  FH_HIP_TRY(hipMemcpyAsync(&count, d_count, sizeof(count), hipMemcpyDeviceToHost, stream));
  FH_HIP_TRY(hipStreamSynchronize(stream));
  if (count < kMemoMin) return FH_OK;
  fh_status st;
  DeviceArray<double> keys, uin;
  DeviceArray<unsigned long long> hash, shash;
  DeviceArray<unsigned> idx, sidx, head, run, uq, ucount;
  DeviceArray<int8_t> umethod;
  DeviceArray<char> temp;
  if ((st = keys.allocate(static_cast<size_t>(count) * 9)) || (st = hash.allocate(count)) ||
      (st = shash.allocate(count)) || (st = idx.allocate(count)) || (st = sidx.allocate(count)) ||
      (st = head.allocate(count)) || (st = run.allocate(count)) || (st = uq.allocate(count)) ||
      (st = ucount.allocate(1)) || (st = uin.allocate(count)) || (st = umethod.allocate(count)))
    return st;
  const unsigned blocks = (count + 255) / 256;
  memo_key_kernel<<<blocks, 256, 0, stream>>>(d_F, queue, pairs, count, keys.ptr, hash.ptr, idx.ptr);
  FH_HIP_TRY(hipGetLastError());
  size_t sort_bytes = 0, scan_bytes = 0;
  FH_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, hash.ptr, shash.ptr, idx.ptr, sidx.ptr,
                                               static_cast<int>(count), 0, 64, stream));
  FH_HIP_TRY(hipcub::DeviceScan::InclusiveSum(nullptr, scan_bytes, head.ptr, run.ptr, static_cast<int>(count),
                                              stream));
  if ((st = temp.allocate(std::max(sort_bytes, scan_bytes)))) return st;
  FH_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(temp.ptr, sort_bytes, hash.ptr, shash.ptr, idx.ptr, sidx.ptr,
                                               static_cast<int>(count), 0, 64, stream));
  memo_head_kernel<<<blocks, 256, 0, stream>>>(shash.ptr, sidx.ptr, keys.ptr, count, head.ptr);
  FH_HIP_TRY(hipGetLastError());
  size_t sb = std::max(sort_bytes, scan_bytes);
  FH_HIP_TRY(hipcub::DeviceScan::InclusiveSum(temp.ptr, sb, head.ptr, run.ptr, static_cast<int>(count), stream));
  memo_compact_kernel<<<blocks, 256, 0, stream>>>(head.ptr, run.ptr, sidx.ptr, count, uq.ptr, ucount.ptr);
  FH_HIP_TRY(hipGetLastError());
  int dev = 0, cus = 256;
  FH_HIP_TRY(hipGetDevice(&dev));
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  memo_eval_kernel<<<static_cast<unsigned>(cus) * 16, 64, 0, stream>>>(d_F, queue, pairs, uq.ptr, ucount.ptr,
                                                                       uin.ptr, umethod.ptr);
  FH_HIP_TRY(hipGetLastError());
  memo_scatter_kernel<<<blocks, 256, 0, stream>>>(d_F, n, queue, pairs, shash.ptr, sidx.ptr, run.ptr, count,
                                                  uin.ptr, umethod.ptr, d_L, d_methods, d_out, err);
  FH_HIP_TRY(hipGetLastError());
  FH_HIP_TRY(hipStreamSynchronize(stream));  // the temporaries are freed on return
  used = true;
  return FH_OK;
}

}  // namespace

fh_status fill_inductance(const fh_filament *d_F, size_t n, double *d_L, int8_t *d_methods,
                          hipStream_t stream, const DevicePairMask *pair_mask) {
  if (n == 0) return FH_OK;
  if (n >= (1u << 30)) return set_error(FH_E_INVALID_ARGUMENT, "too many filaments");
  fh_status st = upload_tables();
  if (st) return st;
  const int nt = static_cast<int>((n + kTile - 1) / kTile);
  const long long tiles = static_cast<long long>(nt) * (nt + 1) / 2;
  const size_t pairs = n * (n + 1) / 2;
  DeviceArray<uint2> queue;
  DeviceArray<unsigned int> qcount;
  DeviceArray<unsigned long long> err;
  if ((st = queue.allocate(pairs))) return st;
  if ((st = qcount.allocate(1))) return st;
  if ((st = err.allocate(1))) return st;
  FH_HIP_TRY(hipMemsetAsync(qcount.ptr, 0, sizeof(unsigned int), stream));
  FH_HIP_TRY(hipMemsetAsync(err.ptr, 0xff, sizeof(unsigned long long), stream));
  if (tiles > 0x7fffffffLL) return set_error(FH_E_INVALID_ARGUMENT, "too many tiles");
  MaskArgs mask{nullptr, 0, 0, nullptr};
  if (pair_mask && pair_mask->group && pair_mask->table)
    mask = MaskArgs{pair_mask->group, static_cast<int>(std::min<size_t>(pair_mask->n, n)),
                    static_cast<int>(pair_mask->groups), pair_mask->table};
  fill_tile_kernel<<<static_cast<unsigned>(tiles), 256, 0, stream>>>(
      d_F, static_cast<int>(n), nt, d_L, d_methods, queue.ptr, qcount.ptr, err.ptr, mask);
  FH_HIP_TRY(hipGetLastError());
  int dev = 0, cus = 256;
  FH_HIP_TRY(hipGetDevice(&dev));
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  bool memo = false;
  if ((st = memo_aligned(d_F, static_cast<int>(n), queue.ptr, nullptr, qcount.ptr, d_L, d_methods, nullptr,
                         err.ptr, stream, memo)))
    return st;
  fill_deferred_kernel<<<static_cast<unsigned>(cus) * 16, 64, 0, stream>>>(
      d_F, static_cast<int>(n), d_L, d_methods, queue.ptr, qcount.ptr, err.ptr, memo ? 1 : 0);
  FH_HIP_TRY(hipGetLastError());
  unsigned long long key = 0;
  FH_HIP_TRY(hipMemcpyAsync(&key, err.ptr, sizeof(key), hipMemcpyDeviceToHost, stream));
  FH_HIP_TRY(hipStreamSynchronize(stream));
  if (key != ~0ULL) {
    const int code = static_cast<int>(key & 3u);
    const unsigned long long pair = key >> 2;
    const int64_t i = static_cast<int64_t>(pair / n), j = static_cast<int64_t>(pair % n);
    return set_error(FH_E_KERNEL,
                     code == 1 ? "overlapping parallel filaments " + std::to_string(i) + " and " +
                                     std::to_string(j) +
                                     " have misaligned cross-sections and coincident sample lines"
                               : "partial inductance of filaments " + std::to_string(i) + " and " +
                                     std::to_string(j) + " is not finite",
                     i, j, 0.0, code);
  }
  return FH_OK;
}

fh_status fill_pairs(const fh_filament *d_F, size_t n, const uint2 *d_pairs, size_t count,
                     double *d_out, hipStream_t stream) {
  if (count == 0) return FH_OK;
  if (n >= (1u << 30) || count >= (1u << 30))
    return set_error(FH_E_INVALID_ARGUMENT, "too many filaments or pairs");
  fh_status st = upload_tables();
  if (st) return st;
  DeviceArray<uint2> queue;
  DeviceArray<unsigned int> qcount;
  DeviceArray<unsigned long long> err;
  if ((st = queue.allocate(count))) return st;
  if ((st = qcount.allocate(1))) return st;
  if ((st = err.allocate(1))) return st;
  FH_HIP_TRY(hipMemsetAsync(qcount.ptr, 0, sizeof(unsigned int), stream));
  FH_HIP_TRY(hipMemsetAsync(err.ptr, 0xff, sizeof(unsigned long long), stream));
  pairs_kernel<<<static_cast<unsigned>((count + 255) / 256), 256, 0, stream>>>(
      d_F, static_cast<int>(n), d_pairs, count, d_out, queue.ptr, qcount.ptr, err.ptr);
  FH_HIP_TRY(hipGetLastError());
  int dev = 0, cus = 256;
  FH_HIP_TRY(hipGetDevice(&dev));
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  bool memo = false;
  if ((st = memo_aligned(d_F, static_cast<int>(n), queue.ptr, d_pairs, qcount.ptr, nullptr, nullptr, d_out,
                         err.ptr, stream, memo)))
    return st;
  pairs_deferred_kernel<<<static_cast<unsigned>(cus) * 16, 64, 0, stream>>>(
      d_F, static_cast<int>(n), d_pairs, d_out, queue.ptr, qcount.ptr, err.ptr, memo ? 1 : 0);
  FH_HIP_TRY(hipGetLastError());
  unsigned long long key = 0;
  FH_HIP_TRY(hipMemcpyAsync(&key, err.ptr, sizeof(key), hipMemcpyDeviceToHost, stream));
  FH_HIP_TRY(hipStreamSynchronize(stream));
  if (key != ~0ULL) {
    const int code = static_cast<int>(key & 3u);
    const unsigned long long pair = key >> 2;
    const int64_t i = static_cast<int64_t>(pair / n), j = static_cast<int64_t>(pair % n);
    return set_error(FH_E_KERNEL,
                     "partial inductance of filaments " + std::to_string(i) + " and " +
                         std::to_string(j) + (code == 1 ? " is not resolved" : " is not finite"),
                     i, j, 0.0, code);
  }
  return FH_OK;
}

// ---------------------------------------------------------------------------
// The fill without allocation: the deferred-pair queue, its count and the
// error key live in `scratch`, a fixed device array (each GPU holds its own
// copy until the process ends), sized for kNoAllocMaxFilaments filaments
// (every pair of the upper triangle queued).  The memo is not used here (it
// needs sort buffers); the deferred kernel evaluates each queued pair itself,
// which gives the same bits.

}  // namespace fh

namespace {
constexpr size_t kNoAllocMaxFilaments = 4000;
constexpr size_t kNoAllocQueueBytes =
    kNoAllocMaxFilaments * (kNoAllocMaxFilaments + 1) / 2 * sizeof(uint2);
constexpr size_t kNoAllocScratchBytes = kNoAllocQueueBytes + 256;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace fh {
namespace {

__global__ void noalloc_clear_kernel(unsigned int *qcount, unsigned long long *err) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *qcount = 0u;
    *err = ~0ULL;
  }
}

// Device address of `scratch` on the current device, looked up once per
// device (the symbol's host-side name is not a device address).
fh_status noalloc_scratch(unsigned char **out) {
  static std::mutex mu;
  static std::vector<unsigned char *> per_device;
  int dev = 0;
  FH_HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  if (static_cast<int>(per_device.size()) <= dev) per_device.resize(dev + 1, nullptr);
  if (!per_device[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(FH_E_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                     hipGetErrorString(e));
    per_device[dev] = static_cast<unsigned char *>(p);
  }
  *out = per_device[dev];
  return FH_OK;
}

}  // namespace

size_t fill_noalloc_max_filaments() { return kNoAllocMaxFilaments; }

fh_status fill_inductance_noalloc(const fh_filament *d_F, size_t n, double *d_L, int8_t *d_methods,
                                  hipStream_t stream) {
  if (n > kNoAllocMaxFilaments)
    return set_error(FH_E_INVALID_ARGUMENT,
                     "fh_partial_inductance_noalloc: " + std::to_string(n) +
                         " filaments exceed the limit of its fixed scratch buffer (" +
                         std::to_string(kNoAllocMaxFilaments) + "); use fh_partial_inductance",
                     static_cast<int64_t>(n), static_cast<int64_t>(kNoAllocMaxFilaments));
  // set-up on this device: the scratch address and the constant tables
  unsigned char *base = nullptr;
  fh_status st = noalloc_scratch(&base);
  if (st) return st;
  if ((st = upload_tables())) return st;
  if (n == 0) return FH_OK;
  uint2 *queue = reinterpret_cast<uint2 *>(base);
  unsigned int *qcount = reinterpret_cast<unsigned int *>(base + kNoAllocQueueBytes);
  unsigned long long *err = reinterpret_cast<unsigned long long *>(base + kNoAllocQueueBytes + 8);
  const int nt = static_cast<int>((n + kTile - 1) / kTile);
  const long long tiles = static_cast<long long>(nt) * (nt + 1) / 2;
  noalloc_clear_kernel<<<1, 64, 0, stream>>>(qcount, err);
  FH_HIP_TRY(hipGetLastError());
  const MaskArgs mask{nullptr, 0, 0, nullptr};
  fill_tile_kernel<<<static_cast<unsigned>(tiles), 256, 0, stream>>>(
      d_F, static_cast<int>(n), nt, d_L, d_methods, queue, qcount, err, mask);
  FH_HIP_TRY(hipGetLastError());
  int dev = 0, cus = 256;
  FH_HIP_TRY(hipGetDevice(&dev));
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  fill_deferred_kernel<<<static_cast<unsigned>(cus) * 16, 64, 0, stream>>>(
      d_F, static_cast<int>(n), d_L, d_methods, queue, qcount, err, 0);
  FH_HIP_TRY(hipGetLastError());
  unsigned long long key = 0;
  FH_HIP_TRY(hipMemcpyAsync(&key, err, sizeof(key), hipMemcpyDeviceToHost, stream));
  FH_HIP_TRY(hipStreamSynchronize(stream));
  if (key != ~0ULL) {
    const int code = static_cast<int>(key & 3u);
    const unsigned long long pair = key >> 2;
    const int64_t i = static_cast<int64_t>(pair / n), j = static_cast<int64_t>(pair % n);
    return set_error(FH_E_KERNEL,
                     code == 1 ? "overlapping parallel filaments " + std::to_string(i) + " and " +
                                     std::to_string(j) +
                                     " have misaligned cross-sections and coincident sample lines"
                               : "partial inductance of filaments " + std::to_string(i) + " and " +
                                     std::to_string(j) + " is not finite",
                     i, j, 0.0, code);
  }
  return FH_OK;
}

void gauss_rule_host(int n, double *nodes, double *weights) { host_rule(n, nodes, weights); }

fh_status device_math(const double *x, size_t n, double *log_out, double *atan_out) {
  if (n == 0) return FH_OK;
  DeviceArray<double> dx, dl, da;
  fh_status st;
  if ((st = dx.allocate(n)) || (st = dl.allocate(n)) || (st = da.allocate(n))) return st;
  FH_HIP_TRY(hipMemcpy(dx.ptr, x, n * sizeof(double), hipMemcpyHostToDevice));
  device_math_kernel<<<static_cast<unsigned>((n + 255) / 256), 256>>>(dx.ptr, n, dl.ptr, da.ptr);
  FH_HIP_TRY(hipGetLastError());
  FH_HIP_TRY(hipMemcpy(log_out, dl.ptr, n * sizeof(double), hipMemcpyDeviceToHost));
  FH_HIP_TRY(hipMemcpy(atan_out, da.ptr, n * sizeof(double), hipMemcpyDeviceToHost));
  return FH_OK;
}

}  // namespace fh
