// Derivative kernels of the partial-inductance fill, for the impedance
// gradients of fh_grad.
//
// A filament pair's value L_fg = 1e-7 x (integral) comes from one of the
// fill's methods (inductance/mod.rs:303-369): point quadrature, the 64-term
// closed form or the aligned quadrature of aligned bars, sampled filaments.
// Its derivative along a direction of the geometry is computed here in
// forward mode: every quantity carries (value, derivative), the filament
// geometry's derivative being the tangent of the discretization.  All
// branches -- the method, the quadrature orders, the closed form's error
// test, the aligned quadrature's panel subdivision -- are taken on the
// values, i.e. held at the nominal geometry, so the result is the exact
// derivative of the (piecewise smooth) fill at that geometry.  The value
// parts follow the fill's operations one for one, so they reproduce its
// bits (checked by the tests).
//
// Contraction: sum over pairs of (I^a_f I^b_g + I^a_g I^b_f) dL_fg (once for
// f = g) per direction, frequency and port pair.  Thread-per-pair work (the
// cheap point quadrature) is reduced per tile; the wave-per-pair methods are
// queued, sorted by (direction, f, g) and reduced in that order, so the sums
// do not depend on scheduling.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "fh_adjoint.hpp"
#include "fh_math.hpp"
#include "fh_wave.hpp"

namespace fh {
namespace {

constexpr int kMaxRule = 16;
constexpr int kPointMaxOrder = 16;    // neumann.rs:34
constexpr int kSampleMaxOrder = 8;    // neumann.rs:40
constexpr int kCloudCapacity = 256;   // neumann.rs:36
constexpr int kCheapPointPairs = 4096;  // inductance/mod.rs:88
constexpr int kFarOrder = 10;         // aligned.rs:69
constexpr int kCornerOrder = 16;      // aligned.rs:71
constexpr double kSnap = 1e-12;       // aligned.rs:74
constexpr int kTile = 16;
constexpr int kInlineEvals = 1024;
// The aligned quadrature's stack: at most 201 frames are live (the frame at
// index k has depth k, and depth 200 is a leaf).  The first kStackLds live in
// LDS -- with the point cloud they share it with, 16 KB, so that eight
// one-wave workgroups fit in a CU -- and the rest, reached only by
// subdivisions 157 deep, in a per-wave area of global memory.
constexpr int kStackCapacity = 202;
constexpr int kStackLds = 157;
constexpr int kStackOvf = kStackCapacity - kStackLds;
constexpr int kMaxOut = 128;          // K p^2 per contraction pass

__constant__ double a_gx[kMaxRule + 1][kMaxRule];
__constant__ double a_gw[kMaxRule + 1][kMaxRule];
__constant__ double a_point_th[kPointMaxOrder];
__constant__ double a_sample_th[kSampleMaxOrder];

enum Kind : uint32_t { kLightPoint = 0u, kHeavyPoint = 1u, kAlignedKind = 2u, kSampled = 3u };

// ---------------------------------------------------------------------------
// Dual numbers.

struct Dual {
  double v, d;
};
__device__ __forceinline__ Dual cst(double v) { return {v, 0.0}; }
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator-(Dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
__device__ __forceinline__ Dual operator+(double a, Dual b) { return {a + b.v, b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
__device__ __forceinline__ Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
__device__ __forceinline__ Dual operator/(double a, Dual b) {
  const double q = a / b.v;
  return {q, -q * b.d / b.v};
}
__device__ __forceinline__ Dual dsqrt(Dual a) {
  const double r = __builtin_sqrt(a.v);
  return {r, r > 0.0 ? a.d / (2.0 * r) : 0.0};
}
__device__ __forceinline__ Dual dlog(Dual a) { return {pm_log(a.v), a.d / a.v}; }
__device__ __forceinline__ Dual datan(Dual a) { return {pm_atan(a.v), a.d / (1.0 + a.v * a.v)}; }
__device__ __forceinline__ Dual dhypot(Dual a, Dual b) {
  const double r = pm_hypot(a.v, b.v);
  return {r, r > 0.0 ? (a.v * a.d + b.v * b.d) / r : 0.0};
}
__device__ __forceinline__ Dual dfabs(Dual a) { return __builtin_signbit(a.v) ? -a : a; }
__device__ __forceinline__ Dual dfmin(Dual a, Dual b) { return b.v < a.v ? b : a; }
__device__ __forceinline__ Dual dfmax(Dual a, Dual b) { return b.v > a.v ? b : a; }

struct V3 {
  Dual x, y, z;
};
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, Dual s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator/(V3 a, Dual s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ Dual dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ Dual norm(V3 a) { return dsqrt(dot(a, a)); }
__device__ __forceinline__ V3 ld(const double *p, const double *t) {
  return {{p[0], t[0]}, {p[1], t[1]}, {p[2], t[2]}};
}
__device__ __forceinline__ double signum(double x) {
  if (x != x) return x;
  return __builtin_signbit(x) ? -1.0 : 1.0;
}

// ---------------------------------------------------------------------------
// Primal decisions (the fill's plan_pair on the values).

struct P3 {
  double x, y, z;
};
__device__ __forceinline__ P3 operator+(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ P3 operator-(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ P3 operator*(P3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ P3 operator/(P3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ double pdot(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ P3 pcross(P3 a, P3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double pnorm(P3 a) { return __builtin_sqrt(pdot(a, a)); }
__device__ __forceinline__ P3 pld(const double *p) { return {p[0], p[1], p[2]}; }

struct PBar {
  P3 centre, axes[3];
  double half[3];
};
__device__ __forceinline__ PBar pbar(const fh_filament &f) {
  PBar b;
  b.centre = (pld(f.start) + pld(f.end)) * 0.5;
  b.axes[0] = pld(f.dir);
  b.axes[1] = pld(f.wdir);
  b.axes[2] = pld(f.hdir);
  b.half[0] = 0.5 * f.length;
  b.half[1] = 0.5 * f.width;
  b.half[2] = 0.5 * f.height;
  return b;
}

__device__ __forceinline__ int order_for(const double *th, int max_order, double half, double distance) {
  if (half <= 0.0) return 1;
  if (distance != distance || distance <= 0.0) return 0;
  const double delta = distance / half;
  for (int i = 0; i < max_order; ++i)
    if (delta >= th[i]) return i + 1;
  return 0;
}

__device__ __forceinline__ double projected_half_extent(const PBar &b, P3 n) {
  double s = 0.0;
  for (int k = 0; k < 3; ++k) s += b.half[k] * __builtin_fabs(pdot(n, b.axes[k]));
  return s;
}

__device__ double separation(const PBar &a, const PBar &b) {
  const P3 delta = b.centre - a.centre;
  const double ra = __builtin_sqrt(a.half[0] * a.half[0] + a.half[1] * a.half[1] + a.half[2] * a.half[2]);
  const double rb = __builtin_sqrt(b.half[0] * b.half[0] + b.half[1] * b.half[1] + b.half[2] * b.half[2]);
  const double distance = pnorm(delta), reach = ra + rb;
  if (distance > 8.0 * reach) return distance - reach;
  double best = -__builtin_inf();
  auto consider = [&](P3 n) {
    const double gap =
        __builtin_fabs(pdot(n, delta)) - projected_half_extent(a, n) - projected_half_extent(b, n);
    best = __builtin_fmax(best, gap);
  };
  for (int k = 0; k < 3; ++k) {
    consider(a.axes[k]);
    consider(b.axes[k]);
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const P3 n = pcross(a.axes[i], b.axes[j]);
      const double nn = pnorm(n);
      if (nn > 1e-9) consider(n / nn);
    }
  return best;
}

__device__ __forceinline__ bool point_orders(const PBar &b, double gap, int o[3]) {
  for (int k = 0; k < 3; ++k) {
    o[k] = order_for(a_point_th, kPointMaxOrder, b.half[k], gap);
    if (!o[k]) return false;
  }
  return o[0] * o[1] * o[2] <= kCloudCapacity;
}

__device__ __forceinline__ long long total_key(double x) {
  long long a = __double_as_longlong(x);
  return a ^ static_cast<long long>(static_cast<unsigned long long>(a >> 63) >> 1);
}

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

// ---------------------------------------------------------------------------
// Dual geometry.

struct Bar {
  V3 centre, axes[3];
  Dual half[3];
};
__device__ __forceinline__ Bar make_bar(const fh_filament &f, const fh_filament &t) {
  Bar b;
  b.centre = (ld(f.start, t.start) + ld(f.end, t.end)) * 0.5;
  b.axes[0] = ld(f.dir, t.dir);
  b.axes[1] = ld(f.wdir, t.wdir);
  b.axes[2] = ld(f.hdir, t.hdir);
  b.half[0] = 0.5 * Dual{f.length, t.length};
  b.half[1] = 0.5 * Dual{f.width, t.width};
  b.half[2] = 0.5 * Dual{f.height, t.height};
  return b;
}

struct Aligned {
  Dual length[2], width[2], height[2], offset[3];
};

// inductance/mod.rs:280-300 (the fill's aligned_of), on duals.
__device__ bool aligned_of(const Bar &a, const Bar &b, Aligned &o) {
  const double kAligned = 1.0 - 1e-12;
  // values, not a choice between addresses (which would put the bar in
  // scratch memory)
  const Dual h1 = b.half[1], h2 = b.half[2];
  const bool same = __builtin_fabs(dot(a.axes[1], b.axes[1]).v) > kAligned;
  if (!same && !(__builtin_fabs(dot(a.axes[1], b.axes[2]).v) > kAligned)) return false;
  const Dual wb{same ? h1.v : h2.v, same ? h1.d : h2.d}, hb{same ? h2.v : h1.v, same ? h2.d : h1.d};
  const V3 delta = b.centre - a.centre;
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

struct Plan {
  Bar a, b;
  Dual cos, area_a, area_b;
  double gap;
  bool parallel, aligned, point, orthogonal;
  // the direction turns the pair's axes relative to each other while the
  // method assumes them fixed (perpendicular: value 0; parallel: the aligned
  // or parallel-line formulas), so its derivative is not available
  bool unsupported;
  int oa[3], ob[3];
  Aligned ab;
};

__device__ __forceinline__ double tnorm(const double *t) {
  return __builtin_sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
}

// plan_pair of the fill: decisions on the values, geometry carried as duals.
__device__ void plan_pair(const fh_filament &fa_in, const fh_filament &ta_in, const fh_filament &fb_in,
                          const fh_filament &tb_in, Plan &p) {
  const fh_filament *fa = &fa_in, *fb = &fb_in, *ta = &ta_in, *tb = &tb_in;
  if (canonical(*fa, *fb) > 0) {
    const fh_filament *t = fa;
    fa = fb;
    fb = t;
    t = ta;
    ta = tb;
    tb = t;
  }
  const V3 da = ld(fa->dir, ta->dir), db = ld(fb->dir, tb->dir);
  p.cos = dot(da, db);
  p.orthogonal = __builtin_fabs(p.cos.v) < 1e-14;
  // how fast the frames turn (0 up to rounding for directions that only
  // move and resize), against the turning rate the end motions could give
  const double turn = tnorm(ta->dir) + tnorm(tb->dir) + tnorm(ta->wdir) + tnorm(tb->wdir) +
                      tnorm(ta->hdir) + tnorm(tb->hdir);
  const double motion = (tnorm(ta->start) + tnorm(ta->end)) / fa->length +
                        (tnorm(tb->start) + tnorm(tb->end)) / fb->length;
  const double turn_tol = 1e-9 * (turn + motion);
  p.unsupported = false;
  if (p.orthogonal) {
    p.unsupported = __builtin_fabs(p.cos.d) > turn_tol;
    return;
  }
  p.a = make_bar(*fa, *ta);
  p.b = make_bar(*fb, *tb);
  p.area_a = Dual{fa->width, ta->width} * Dual{fa->height, ta->height};
  p.area_b = Dual{fb->width, tb->width} * Dual{fb->height, tb->height};
  const P3 pda = pld(fa->dir), pdb = pld(fb->dir);
  p.parallel = pnorm(pcross(pda, pdb)) < 1e-6;
  p.aligned = p.parallel ? aligned_of(p.a, p.b, p.ab) : false;
  const PBar pa = pbar(*fa), pb = pbar(*fb);
  p.gap = separation(pa, pb);
  bool have = point_orders(pa, p.gap, p.oa);
  if (have) have = point_orders(pb, p.gap, p.ob);
  if (have && p.aligned) {
    const int pairs = p.oa[0] * p.oa[1] * p.oa[2] * (p.ob[0] * p.ob[1] * p.ob[2]);
    if (pairs > kCheapPointPairs) have = false;
  }
  p.point = have;
  if (turn > 0.0 && p.parallel && !p.point) {
    const V3 c = cross(da, db);
    bool rel = __builtin_sqrt(c.x.d * c.x.d + c.y.d * c.y.d + c.z.d * c.z.d) > turn_tol;
    if (p.aligned) {
      // the cross-section axis of b not matched to a's width axis stays
      // perpendicular to it
      const V3 wa = ld(fa->wdir, ta->wdir), wb = ld(fb->wdir, tb->wdir), hb = ld(fb->hdir, tb->hdir);
      const V3 other = __builtin_fabs(dot(wa, wb).v) > 1.0 - 1e-12 ? hb : wb;
      rel = rel || __builtin_fabs(dot(wa, other).d) > turn_tol;
    }
    p.unsupported = rel;
  }
}

__device__ __forceinline__ void rule_on(int n, int i, Dual a, Dual b, Dual &x, Dual &w) {
  const Dual mid = 0.5 * (a + b), half = 0.5 * (b - a);
  x = mid + half * a_gx[n][i];
  w = half * a_gw[n][i];
}
__device__ __forceinline__ void rule_on(int n, int i, double a, double b, double &x, double &w) {
  const double mid = 0.5 * (a + b), half = 0.5 * (b - a);
  x = mid + half * a_gx[n][i];
  w = half * a_gw[n][i];
}

// ---------------------------------------------------------------------------
// Point quadrature (neumann.rs:160-266).

// The four lane sums rotate through r0..r3 (the next term always goes to
// r0); r_j holds lane (k + j) mod 4 after k terms, undone at the end.
__device__ Dual inner_sum(V3 pa, const Bar &b, const int ob[3]) {
  Dual r0 = cst(0.0), r1 = cst(0.0), r2 = cst(0.0), r3 = cst(0.0);
  int k = 0;
  const Dual s1 = 2.0 * b.half[1], s2 = 2.0 * b.half[2];
  for (int il = 0; il < ob[0]; ++il) {
    Dual xl, wl;
    rule_on(ob[0], il, -b.half[0], b.half[0], xl, wl);
    const V3 p1 = b.centre + b.axes[0] * xl;
    for (int iw = 0; iw < ob[1]; ++iw) {
      double xw, ww;
      rule_on(ob[1], iw, -0.5, 0.5, xw, ww);
      const V3 p2 = p1 + b.axes[1] * (s1 * xw);
      const Dual wlw = wl * ww;
      for (int ih = 0; ih < ob[2]; ++ih) {
        double xh, wh;
        rule_on(ob[2], ih, -0.5, 0.5, xh, wh);
        const V3 p3 = p2 + b.axes[2] * (s2 * xh);
        const Dual w = wlw * wh;
        const Dual dx = pa.x - p3.x, dy = pa.y - p3.y, dz = pa.z - p3.z;
        const Dual t = w / dsqrt(dx * dx + dy * dy + dz * dz);
        const Dual s = r0 + t;
        r0 = r1;
        r1 = r2;
        r2 = r3;
        r3 = s;
        ++k;
      }
    }
  }
  const int n = k & 3;
  const Dual in0 = n == 0 ? r0 : n == 1 ? r3 : n == 2 ? r2 : r1;
  const Dual in1 = n == 0 ? r1 : n == 1 ? r0 : n == 2 ? r3 : r2;
  const Dual in2 = n == 0 ? r2 : n == 1 ? r1 : n == 2 ? r0 : r3;
  const Dual in3 = n == 0 ? r3 : n == 1 ? r2 : n == 2 ? r1 : r0;
  return (in0 + in1) + (in2 + in3);
}

__device__ __forceinline__ void cloud_point(const Bar &a, const int oa[3], int index, V3 &p, Dual &w) {
  const int ih = index % oa[2];
  const int iw = (index / oa[2]) % oa[1];
  const int il = index / (oa[2] * oa[1]);
  Dual xl, wl;
  double xw, ww, xh, wh;
  rule_on(oa[0], il, -a.half[0], a.half[0], xl, wl);
  rule_on(oa[1], iw, -0.5, 0.5, xw, ww);
  rule_on(oa[2], ih, -0.5, 0.5, xh, wh);
  p = ((a.centre + a.axes[0] * xl) + a.axes[1] * (2.0 * a.half[1] * xw)) + a.axes[2] * (2.0 * a.half[2] * xh);
  w = wl * ww * wh;
}

__device__ Dual point_sum_serial(const Bar &a, const int oa[3], const Bar &b, const int ob[3]) {
  Dual total = cst(0.0);
  const Dual s1 = 2.0 * a.half[1], s2 = 2.0 * a.half[2];
  for (int il = 0; il < oa[0]; ++il) {
    Dual xl, wl;
    rule_on(oa[0], il, -a.half[0], a.half[0], xl, wl);
    const V3 p1 = a.centre + a.axes[0] * xl;
    for (int iw = 0; iw < oa[1]; ++iw) {
      double xw, ww;
      rule_on(oa[1], iw, -0.5, 0.5, xw, ww);
      const V3 p2 = p1 + a.axes[1] * (s1 * xw);
      const Dual wlw = wl * ww;
      for (int ih = 0; ih < oa[2]; ++ih) {
        double xh, wh;
        rule_on(oa[2], ih, -0.5, 0.5, xh, wh);
        const V3 p3 = p2 + a.axes[2] * (s2 * xh);
        total = total + (wlw * wh) * inner_sum(p3, b, ob);
      }
    }
  }
  return total;
}

// ---------------------------------------------------------------------------
// bar.rs closed form.  A term c x log(.) with x = 0 keeps its derivative
// c dx log(r), the limit of its x-derivative.

__device__ __forceinline__ Dual log_term(Dual c, Dual x, Dual r, Dual rest_sq) {
  if (c.v == 0.0) return cst(0.0);
  if (x.v == 0.0) return {0.0, r.v > 0.0 ? c.v * x.d * pm_log(r.v) : 0.0};
  const Dual argument = x.v >= 0.0 ? x + r : rest_sq / (r - x);
  return c * x * dlog(argument);
}

__device__ __forceinline__ Dual atan_term(Dual x, Dual y, Dual z, Dual r) {
  const Dual product = x * y * z;
  if (product.v == 0.0) return cst(0.0);
  return -(product * z * z / 6.0) * datan(x * y / (z * r));
}

__device__ Dual primitive(Dual x, Dual y, Dual z) {
  const Dual x2 = x * x, y2 = y * y, z2 = z * z;
  const Dual x4 = x2 * x2, y4 = y2 * y2, z4 = z2 * z2;
  const Dual r = dsqrt(x2 + y2 + z2);
  return log_term(y2 * z2 / 4.0 - y4 / 24.0 - z4 / 24.0, x, r, y2 + z2) +
         log_term(x2 * z2 / 4.0 - x4 / 24.0 - z4 / 24.0, y, r, x2 + z2) +
         log_term(x2 * y2 / 4.0 - x4 / 24.0 - y4 / 24.0, z, r, x2 + y2) +
         (x4 + y4 + z4 - 3.0 * (x2 * y2 + y2 * z2 + z2 * x2)) * r / 60.0 + atan_term(x, y, z, r) +
         atan_term(x, z, y, r) + atan_term(y, z, x, r);
}

// Corner coordinate c (the fill's corner, on duals), chosen by value.
__device__ __forceinline__ Dual corner(Dual e1, Dual e2, Dual off, int c) {
  const Dual sum = 0.5 * (e1 + e2), diff = 0.5 * (e1 - e2);
  const bool outer = c == 0 || c == 3;
  const Dual t{outer ? sum.v : diff.v, outer ? sum.d : diff.d};
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

// lines.rs parallel lines.  The value is the fill's arithmetic.  Its terms
// |q| log(|q| + r) - r have kinks at q = 0 (touching ends) that cancel
// against the overlap term: with sum_i s_i |q_i| = 2 overlap the result is
// sum_i s_i (q_i asinh(q_i / rho) - r_i), smooth for rho > 0, whose
// derivative sum_i s_i (asinh(q_i / rho) dq_i - r_i / rho drho) is taken
// here.  For rho = 0 (collinear, apart) the terms are |q| log(2 |q|) - |q|
// and the rho-derivative vanishes.
__device__ __forceinline__ bool lines_parallel(Dual l1, Dual l2, Dual offset, Dual rho, Dual &out) {
  const Dual q[4] = {offset + l2, offset, offset + l2 - l1, offset - l1};
  const double sgn[4] = {1.0, -1.0, -1.0, 1.0};
  const double rho_sq = rho.v * rho.v;
  const double smooth = parallel_primitive(q[0].v, rho_sq) - parallel_primitive(q[1].v, rho_sq) -
                        parallel_primitive(q[2].v, rho_sq) + parallel_primitive(q[3].v, rho_sq);
  const double overlap =
      __builtin_fmax(__builtin_fmin(l1.v, q[0].v) - __builtin_fmax(offset.v, 0.0), 0.0);
  double value;
  if (overlap == 0.0) value = smooth;
  else if (rho.v > 0.0) value = smooth - 2.0 * overlap * pm_log(rho.v);
  else return false;
  double d = 0.0;
  if (rho.v > 0.0) {
    const double log_rho = pm_log(rho.v);
    for (int i = 0; i < 4; ++i) {
      const double a = __builtin_fabs(q[i].v), r = __builtin_sqrt(a * a + rho_sq);
      const double as = pm_log(a + r) - log_rho;  // asinh(|q| / rho)
      d += sgn[i] * ((q[i].v < 0.0 ? -as : as) * q[i].d - (r / rho.v) * rho.d);
    }
  } else {
    for (int i = 0; i < 4; ++i)
      if (q[i].v != 0.0)
        d += sgn[i] * (q[i].v < 0.0 ? -1.0 : 1.0) * pm_log(2.0 * __builtin_fabs(q[i].v)) * q[i].d;
  }
  out = Dual{value, d};
  return true;
}

// x log(argument); at x = 0 the derivative keeps dx log(argument).
__device__ __forceinline__ Dual skew_log_term(Dual x, Dual big_r, Dual a, Dual res) {
  const Dual argument = a.v >= 0.0 ? big_r + a : res / (big_r - a);
  if (!(argument.v > 0.0)) return cst(0.0);
  if (x.v == 0.0) return Dual{0.0, x.d * pm_log(argument.v)};
  return x * dlog(argument);
}

__device__ Dual skew_primitive(Dual s, Dual t, Dual d, Dual c, Dual sn) {
  const Dual sin_sq = sn * sn;
  const Dual versine = sin_sq / (1.0 + dfabs(c));
  Dual planar_sq, a_s, a_t;
  if (c.v >= 0.0) {
    const Dual diff = s - t;
    planar_sq = diff * diff + 2.0 * s * t * versine;
    a_s = s * versine - diff;
    a_t = t * versine + diff;
  } else {
    const Dual sum = s + t;
    planar_sq = sum * sum - 2.0 * s * t * versine;
    a_s = sum - s * versine;
    a_t = sum - t * versine;
  }
  const Dual big_r = dsqrt(dfmax(planar_sq + d * d, cst(0.0)));
  const Dual logs = skew_log_term(s, big_r, a_s, s * s * sin_sq + d * d) +
                    skew_log_term(t, big_r, a_t, t * t * sin_sq + d * d);
  if (d.v == 0.0 || big_r.v == 0.0) return logs;
  return logs - (d / sn) * datan((d * d * c + s * t * sin_sq) / (d * big_r * sn));
}

__device__ bool lines_skew(V3 a0, V3 a1, V3 b0, V3 b1, double sin_min, Dual &out) {
  const Dual la = norm(a1 - a0), lb = norm(b1 - b0);
  const V3 u = (a1 - a0) / la, v = (b1 - b0) / lb;
  const Dual c = dot(u, v);
  const Dual sn = norm(cross(u, v));
  if (sn.v != sn.v || sn.v < sin_min) return false;
  const V3 w0 = a0 - b0;
  const Dual du = dot(u, w0), dv = dot(v, w0);
  const Dual sin_sq = sn * sn;
  const Dual s0 = (c * dv - du) / sin_sq;
  const Dual t0 = (dv - c * du) / sin_sq;
  Dual d = norm((w0 + u * s0) - v * t0);
  if (d.v <= 1e-14 * (la.v + lb.v + norm(w0).v)) d = cst(0.0);
  const Dual s_lo = -s0, s_hi = la - s0, t_lo = -t0, t_hi = lb - t0;
  out = skew_primitive(s_hi, t_hi, d, c, sn) - skew_primitive(s_hi, t_lo, d, c, sn) -
        skew_primitive(s_lo, t_hi, d, c, sn) + skew_primitive(s_lo, t_lo, d, c, sn);
  return true;
}

// ---------------------------------------------------------------------------
// aligned.rs

__device__ __forceinline__ Dual overlap_weight(Dual e1, Dual e2, Dual off, Dual u) {
  const Dual hi = dfmin(0.5 * e1, off + 0.5 * e2 - u);
  const Dual lo = dfmax(-0.5 * e1, off - 0.5 * e2 - u);
  return dfmax(hi - lo, cst(0.0));
}

__device__ __forceinline__ Dual integrand(const Aligned &b, Dual axial_start, Dual u, Dual v) {
  const Dual weight = overlap_weight(b.width[0], b.width[1], b.offset[1], u) *
                      overlap_weight(b.height[0], b.height[1], b.offset[2], v);
  if (weight.v == 0.0) return cst(0.0);
  Dual kernel;
  if (!lines_parallel(b.length[0], b.length[1], axial_start, dhypot(u, v), kernel)) kernel = cst(0.0);
  return weight * kernel;
}

__device__ int break_points(Dual e1, Dual e2, Dual off, Dual *out) {
  const Dual sum = 0.5 * (e1 + e2), diff = 0.5 * dfabs(e1 - e2);
  Dual p[5];
  int n = 4;
  const Dual raw[4] = {off - sum, off - diff, off + diff, off + sum};
  for (int i = 0; i < 4; ++i) p[i] = __builtin_fabs(raw[i].v) <= kSnap * sum.v ? cst(0.0) : raw[i];
  if (p[0].v < 0.0 && p[3].v > 0.0) p[n++] = cst(0.0);
  for (int i = 1; i < n; ++i) {
    const Dual v = p[i];
    int j = i - 1;
    while (j >= 0 && total_key(p[j].v) > total_key(v.v)) {
      p[j + 1] = p[j];
      --j;
    }
    p[j + 1] = v;
  }
  int m = 0;
  for (int i = 0; i < n; ++i)
    if (m == 0 || !(out[m - 1].v == p[i].v)) out[m++] = p[i];
  return m;
}

__device__ __forceinline__ Dual wave_ordered_sum(Dual acc, Dual term, int count, Dual *lds) {
  __syncthreads();
  lds[threadIdx.x] = term;
  __syncthreads();
  for (int q = 0; q < count; ++q) acc = acc + lds[q];
  return acc;
}

__device__ Dual direct_wave(const Aligned &b, Dual axial_start, Dual u0, Dual u1, Dual v0, Dual v1, Dual *lds) {
  Dual sum = cst(0.0);
  for (int base = 0; base < kFarOrder * kFarOrder; base += 64) {
    const int q = base + static_cast<int>(threadIdx.x);
    Dual term = cst(0.0);
    if (q < kFarOrder * kFarOrder) {
      Dual uu, wu, vv, wv;
      rule_on(kFarOrder, q / kFarOrder, u0, u1, uu, wu);
      rule_on(kFarOrder, q % kFarOrder, v0, v1, vv, wv);
      term = wu * wv * integrand(b, axial_start, uu, vv);
    }
    const int count = kFarOrder * kFarOrder - base < 64 ? kFarOrder * kFarOrder - base : 64;
    sum = wave_ordered_sum(sum, term, count, lds);
  }
  return sum;
}

__device__ Dual corner_wave(const Aligned &b, Dual axial_start, Dual a, Dual bb, Dual *lds) {
  Dual sum = cst(0.0);
  for (int base = 0; base < kCornerOrder * kCornerOrder; base += 64) {
    const int q = base + static_cast<int>(threadIdx.x);
    double tau, w_tau, t, w_t;
    rule_on(kCornerOrder, q / kCornerOrder, 0.0, 1.0, tau, w_tau);
    rule_on(kCornerOrder, q % kCornerOrder, 0.0, 1.0, t, w_t);
    const double s = tau * tau * tau;
    const double radial = w_tau * 3.0 * tau * tau * s;
    const Dual value = integrand(b, axial_start, a * s, bb * s * t) + integrand(b, axial_start, a * s * t, bb * s);
    const Dual term = radial * w_t * value;
    sum = wave_ordered_sum(sum, term, 64, lds);
  }
  return sum * dfabs(a * bb);
}

struct Frame {
  Dual u0, u1, v0, v1, at, acc;
  int16_t depth;
  int8_t stage, axis;
};
static_assert(sizeof(Frame) == 104, "Frame layout");

// LDS of a deferred wavefront: the bottom of the aligned quadrature's stack,
// or the inner cloud (x, y, z, w) of a heavy point quadrature; one pair uses
// one of them.
union DeferredScratch {
  Frame stack[kStackLds];
  Dual cloud[4 * kCloudCapacity];
};
static_assert(sizeof(DeferredScratch) == 16384, "16 KB of LDS");

// The wave's stack: frames below kStackLds in LDS, the others in its global
// overflow area (every lane reads and writes the same values).
struct FrameStack {
  Frame *lds, *ovf;
  __device__ __forceinline__ Frame get(int i) const {
    if (i < kStackLds) return lds[i];
    return ovf[i - kStackLds];
  }
  __device__ __forceinline__ void put(int i, const Frame &f) const {
    if (i < kStackLds) lds[i] = f;
    else ovf[i - kStackLds] = f;
  }
};

__device__ __forceinline__ Dual shfl_xor_dual(Dual a, int m) {
  return {__shfl_xor(a.v, m), __shfl_xor(a.d, m)};
}

__device__ __forceinline__ double gap_of(double r0, double r1) {
  if (r0 > 0.0) return r0;
  if (r1 < 0.0) return -r1;
  return 0.0;
}

// aligned.rs:153-199 with the explicit stack of the fill; the subdivision is
// decided on the values, the panel limits carry their derivatives.
__device__ Dual panel_wave(const Aligned &b, Dual axial_start, Dual axial_scale, Dual u0, Dual u1, Dual v0,
                           Dual v1, const FrameStack &stack, Dual *lds) {
  int sp = 0;
  stack.put(0, Frame{u0, u1, v0, v1, cst(0.0), cst(0.0), 0, 0, 0});
  sp = 1;
  Dual ret = cst(0.0);
  bool have = false;
  for (;;) {
    if (!have) {
      const Frame f = stack.get(sp - 1);
      const Dual su = f.u1 - f.u0, sv = f.v1 - f.v0;
      int axis = -1;
      Dual at = cst(0.0);
      Dual leaf = cst(0.0);
      bool is_leaf = true;
      if (su.v <= 0.0 || sv.v <= 0.0) {
        leaf = cst(0.0);
      } else {
        const double distance = pm_hypot(gap_of(f.u0.v, f.u1.v), gap_of(f.v0.v, f.v1.v));
        const double size = __builtin_fmax(su.v, sv.v);
        if (f.depth >= 200) {
          leaf = direct_wave(b, axial_start, f.u0, f.u1, f.v0, f.v1, lds);
        } else if (distance > 0.0) {
          if (size <= distance) {
            leaf = direct_wave(b, axial_start, f.u0, f.u1, f.v0, f.v1, lds);
          } else if (su.v >= sv.v) {
            is_leaf = false;
            axis = 0;
            at = 0.5 * (f.u0 + f.u1);
          } else {
            is_leaf = false;
            axis = 1;
            at = 0.5 * (f.v0 + f.v1);
          }
        } else {
          const Dual target = dfmin(dfmin(su, sv), axial_scale);
          if (su.v > 2.0 * target.v) {
            is_leaf = false;
            axis = 0;
            at = f.u0.v == 0.0 ? target : -target;
          } else if (sv.v > 2.0 * target.v) {
            is_leaf = false;
            axis = 1;
            at = f.v0.v == 0.0 ? target : -target;
          } else {
            leaf = corner_wave(b, axial_start, f.u0.v == 0.0 ? f.u1 : f.u0, f.v0.v == 0.0 ? f.v1 : f.v0, lds);
          }
        }
      }
      if (is_leaf || sp >= kStackCapacity) {
        if (!is_leaf) leaf = direct_wave(b, axial_start, f.u0, f.u1, f.v0, f.v1, lds);
        ret = leaf;
        have = true;
        --sp;
      } else {
        Frame t = f;
        t.stage = 1;
        t.axis = static_cast<int8_t>(axis);
        t.at = at;
        stack.put(sp - 1, t);
        Frame c = f;
        c.depth = static_cast<int16_t>(f.depth + 1);
        c.stage = 0;
        if (axis == 0) c.u1 = at;
        else c.v1 = at;
        stack.put(sp++, c);
      }
    } else {
      if (sp == 0) return ret;
      Frame p = stack.get(sp - 1);
      if (p.stage == 1) {
        p.acc = ret;
        p.stage = 2;
        stack.put(sp - 1, p);
        Frame c = p;
        c.depth = static_cast<int16_t>(p.depth + 1);
        c.stage = 0;
        if (p.axis == 0) c.u0 = p.at;
        else c.v0 = p.at;
        stack.put(sp++, c);
        have = false;
      } else {
        ret = p.acc + ret;
        --sp;
      }
    }
  }
}

__device__ Dual aligned_quadrature_wave(const Aligned &b, const FrameStack &stack, Dual *lds) {
  const Dual axial_start = b.offset[0] + 0.5 * (b.length[0] - b.length[1]);
  const Dual l1 = b.length[0], l2 = b.length[1];
  const Dual sep[4] = {axial_start + l2, axial_start, axial_start + l2 - l1, axial_start - l1};
  double largest = 0.0;
  for (int i = 0; i < 4; ++i) largest = __builtin_fmax(largest, __builtin_fabs(sep[i].v));
  Dual axial_scale = cst(__builtin_inf());
  for (int i = 0; i < 4; ++i) {
    const Dual q = dfabs(sep[i]);
    if (q.v > 1e-9 * largest) axial_scale = dfmin(axial_scale, q);
  }
  Dual us[5], vs[5];
  const int nu = break_points(b.width[0], b.width[1], b.offset[1], us);
  const int nv = break_points(b.height[0], b.height[1], b.offset[2], vs);
  Dual sum = cst(0.0);
  for (int i = 0; i + 1 < nu; ++i)
    for (int j = 0; j + 1 < nv; ++j)
      sum = sum + panel_wave(b, axial_start, axial_scale, us[i], us[i + 1], vs[j], vs[j + 1], stack, lds);
  return sum;
}

__device__ Dual aligned_dispatch_wave(const Aligned &b, const FrameStack &stack, Dual *lds) {
  const int lane = static_cast<int>(threadIdx.x);
  const int i = lane >> 4, j = (lane >> 2) & 3, k = lane & 3;
  const Dual qx = corner(b.length[0], b.length[1], b.offset[0], i);
  const Dual qy = corner(b.width[0], b.width[1], b.offset[1], j);
  const Dual qz = corner(b.height[0], b.height[1], b.offset[2], k);
  const Dual term = corner_sign(i) * corner_sign(j) * corner_sign(k) * primitive(qx, qy, qz);
  __syncthreads();
  lds[lane] = term;
  __syncthreads();
  Dual value = cst(0.0);
  double magnitude = 0.0;
  for (int q = 0; q < 64; ++q) {
    value = value + lds[q];
    magnitude += __builtin_fabs(lds[q].v);
  }
  const double err = value.v > 0.0 ? 8.0 * 2.220446049250313e-16 * magnitude / value.v : __builtin_inf();
  if (err <= 1e-10) return value;
  return aligned_quadrature_wave(b, stack, lds);
}

// The fill's deferred evaluation (one wave), on duals: 1e-7 x integral.
__device__ __forceinline__ Dual deferred_eval(const Plan &p, uint32_t kind, Dual *lds, DeferredScratch &sh,
                                              Frame *ovf, int lane, bool &ok) {
  Dual integral = cst(0.0);
  ok = true;
  if (kind == kHeavyPoint) {
    // As the fill's: b's cloud formed once in LDS, four lanes per outer
    // point i each running one of the four interleaved lane sums, combined
    // as (l0 + l1) + (l2 + l3); the outer sum in order of i.
    const int na = p.oa[0] * p.oa[1] * p.oa[2], nb = p.ob[0] * p.ob[1] * p.ob[2];
    Dual *cx = sh.cloud, *cy = cx + kCloudCapacity, *cz = cy + kCloudCapacity, *cw = cz + kCloudCapacity;
    __syncthreads();
    for (int q = lane; q < nb; q += 64) {
      V3 pb;
      Dual wb;
      cloud_point(p.b, p.ob, q, pb, wb);
      cx[q] = pb.x;
      cy[q] = pb.y;
      cz[q] = pb.z;
      cw[q] = wb;
    }
    __syncthreads();
    for (int base = 0; base < 4 * na; base += 64) {
      const int q = base + lane, i = q >> 2, slot = q & 3;
      Dual acc = cst(0.0), wa = cst(0.0);
      if (i < na) {
        V3 pa;
        cloud_point(p.a, p.oa, i, pa, wa);
        for (int jj = slot; jj < nb; jj += 4) {
          const Dual dx = pa.x - cx[jj], dy = pa.y - cy[jj], dz = pa.z - cz[jj];
          acc = acc + cw[jj] / dsqrt(dx * dx + dy * dy + dz * dz);
        }
      }
      const Dual pair = acc + shfl_xor_dual(acc, 1);
      const Dual inner = pair + shfl_xor_dual(pair, 2);
      if (i < na && slot == 0) lds[i] = wa * inner;
    }
    __syncthreads();
    Dual total = cst(0.0);
    for (int idx = 0; idx < na; ++idx) total = total + lds[idx];
    integral = total * p.cos;
  } else if (kind == kAlignedKind) {
    const Dual in = aligned_dispatch_wave(p.ab, FrameStack{sh.stack, ovf}, lds);
    const Dual scale = signum(p.cos.v) / (p.area_a * p.area_b);
    integral = in * scale;
  } else {
    int oa[2], ob[2];
    const double halves[4] = {p.a.half[1].v, p.a.half[2].v, p.b.half[1].v, p.b.half[2].v};
    int *targets[4] = {&oa[0], &oa[1], &ob[0], &ob[1]};
    for (int t = 0; t < 4; ++t) {
      const int o = order_for(a_sample_th, kSampleMaxOrder, halves[t], p.gap);
      *targets[t] = o ? o : kSampleMaxOrder + (t >= 2 ? 1 : 0);
    }
    const int na = oa[0] * oa[1], nb = ob[0] * ob[1], total_pairs = na * nb;
    const Dual la = 2.0 * p.a.half[0], lb = 2.0 * p.b.half[0];
    Dual value = cst(0.0);
    int bad = 0;
    for (int base = 0; base < total_pairs; base += 64) {
      const int idx = base + lane;
      Dual term = cst(0.0);
      if (idx < total_pairs) {
        const int ia = idx / nb, ib = idx % nb;
        double xw, ww, xh, wh;
        rule_on(oa[0], ia / oa[1], -0.5, 0.5, xw, ww);
        rule_on(oa[1], ia % oa[1], -0.5, 0.5, xh, wh);
        const V3 pa = p.a.centre + (p.a.axes[1] * (2.0 * p.a.half[1] * xw) + p.a.axes[2] * (2.0 * p.a.half[2] * xh));
        const double wa = ww * wh;
        rule_on(ob[0], ib / ob[1], -0.5, 0.5, xw, ww);
        rule_on(ob[1], ib % ob[1], -0.5, 0.5, xh, wh);
        const V3 pb = p.b.centre + (p.b.axes[1] * (2.0 * p.b.half[1] * xw) + p.b.axes[2] * (2.0 * p.b.half[2] * xh));
        const double wb = ww * wh;
        Dual kernel = cst(0.0);
        bool good;
        if (p.parallel) {
          const V3 axis = p.a.axes[0];
          const V3 delta = pb - pa;
          const Dual along = dot(delta, axis);
          const Dual rho = norm(delta - axis * along);
          good = lines_parallel(la, lb, along + 0.5 * (la - lb), rho, kernel);
        } else {
          const V3 ha = p.a.axes[0] * p.a.half[0], hb = p.b.axes[0] * p.b.half[0];
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
    const Dual direction = p.parallel ? cst(signum(p.cos.v)) : p.cos;
    integral = value * direction;
  }
  return 1.0e-7 * integral;
}

// ---------------------------------------------------------------------------
// Contraction.

struct Weights {
  const double2 *I;  // K x p x n
  int n, K, p;
};

// (I^a_f I^b_g + I^a_g I^b_f) for f != g, I^a_f I^b_f for f = g, at output
// o = (k p + a) p + b.
__device__ __forceinline__ double2 weight(const Weights &W, int f, int g, int o) {
  const int b = o % W.p, a = (o / W.p) % W.p, k = o / (W.p * W.p);
  const double2 *Ik = W.I + static_cast<size_t>(k) * W.p * W.n;
  const double2 af = Ik[static_cast<size_t>(a) * W.n + f], bg = Ik[static_cast<size_t>(b) * W.n + g];
  double re = af.x * bg.x - af.y * bg.y, im = af.x * bg.y + af.y * bg.x;
  if (f != g) {
    const double2 ag = Ik[static_cast<size_t>(a) * W.n + g], bf = Ik[static_cast<size_t>(b) * W.n + f];
    re += ag.x * bf.x - ag.y * bf.y;
    im += ag.x * bf.y + ag.y * bf.x;
  }
  return make_double2(re, im);
}

// Upper-triangular tile t -> (ti, tj), ti <= tj, rows of length nt - ti.
__device__ __forceinline__ void tile_coords(long long t, int nt, int &ti, int &tj) {
  const double disc = (2.0 * nt + 1.0) * (2.0 * nt + 1.0) - 8.0 * static_cast<double>(t);
  int r = static_cast<int>(((2.0 * nt + 1.0) - __builtin_sqrt(disc)) * 0.5);
  if (r < 0) r = 0;
  if (r > nt - 1) r = nt - 1;
  auto start = [nt](long long row) { return row * nt - row * (row - 1) / 2; };
  while (r > 0 && start(r) > t) --r;
  while (r + 1 < nt && start(r + 1) <= t) ++r;
  ti = r;
  tj = r + static_cast<int>(t - start(r));
}

// out[o] = block sum of term(o) over the 256 threads, o < nout, in a fixed
// order (lanes by butterfly, then the four waves).
template <typename Term>
__device__ __forceinline__ void block_sum_terms(Term term, int nout, double2 *out) {
  __shared__ double2 red[4][kMaxOut];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int o = 0; o < nout; ++o) {
    const double2 t = term(o);
    double re = t.x, im = t.y;
    for (int d = 32; d > 0; d >>= 1) {
      re += __shfl_xor(re, d);
      im += __shfl_xor(im, d);
    }
    if (lane == 0) red[wave][o] = make_double2(re, im);
  }
  __syncthreads();
  if (static_cast<int>(threadIdx.x) < nout) {
    const int o = threadIdx.x;
    out[o] = make_double2((red[0][o].x + red[1][o].x) + (red[2][o].x + red[3][o].x),
                          (red[0][o].y + red[1][o].y) + (red[2][o].y + red[3][o].y));
  }
}

// moving[q n + f] = 1 when direction q moves filament f (a nonzero tangent).
// A pair of unmoved filaments has a zero derivative: every derivative part
// is linear in the tangents.
__global__ void moving_kernel(const fh_filament *__restrict__ T, size_t total, uint8_t *__restrict__ moving) {
  const size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (e >= total) return;
  const double *t = reinterpret_cast<const double *>(T + e);
  bool any = false;
  for (int w = 0; w < static_cast<int>(sizeof(fh_filament) / sizeof(double)); ++w) any = any || t[w] != 0.0;
  moving[e] = any ? 1 : 0;
}

// One 16 x 16 tile of pairs (f <= g) per workgroup and direction q
// (blockIdx.y).  Light point quadrature is differentiated by the pair's
// thread; the other methods are queued once (by the q = 0 workgroups) as
// (f | kind << 30, g) for one wave per pair and direction.
// Three waves per SIMD (168 registers, some spilled): 462 -> 271 ms on the
// 6156-filament spiral against two; four spill more and measure 400 ms.
__global__ void __launch_bounds__(256, 3) dtile_kernel(const fh_filament *__restrict__ F, int n, int nt,
                                                    const fh_filament *__restrict__ T,
                                                    const uint8_t *__restrict__ moving, Weights W, int nout,
                                                    double2 *__restrict__ partial, uint2 *__restrict__ queue,
                                                    unsigned *__restrict__ qcount, int *__restrict__ flags) {
  int ti, tj;
  tile_coords(blockIdx.x, nt, ti, tj);
  const int q = blockIdx.y;
  const int f = ti * kTile + static_cast<int>(threadIdx.x) / kTile;
  const int g = tj * kTile + static_cast<int>(threadIdx.x) % kTile;
  double d = 0.0;
  bool queued = false;
  uint32_t kind = 0;
  if (f < n && g < n && f <= g) {
    const size_t qn = static_cast<size_t>(q) * n;
    const bool moves = moving[qn + f] || moving[qn + g];
    if (moves || q == 0) {
      Plan p;
      plan_pair(F[f], T[qn + f], F[g], T[qn + g], p);
      if (p.unsupported) atomicOr(flags, 2);
      if (!p.orthogonal) {
        const bool light = p.point && p.oa[0] * p.oa[1] * p.oa[2] * (p.ob[0] * p.ob[1] * p.ob[2]) <= kInlineEvals;
        if (light && !p.aligned) {
          if (moves) d = (1.0e-7 * (point_sum_serial(p.a, p.oa, p.b, p.ob) * p.cos)).d;
        } else if (q == 0) {
          // aligned light point quadratures go through their canonical
          // aligned numbers, with the queued pairs
          kind = light ? kLightPoint : (p.point ? kHeavyPoint : (p.aligned ? kAlignedKind : kSampled));
          queued = true;
        }
      }
    }
  }
  {
    const unsigned slot = wave_append(queued, qcount);
    if (queued) queue[slot] = make_uint2(static_cast<uint32_t>(f) | (kind << 30), static_cast<uint32_t>(g));
  }
  auto term = [&](int o) {
    if (d == 0.0) return make_double2(0.0, 0.0);
    const double2 w = weight(W, f, g, o);
    return make_double2(w.x * d, w.y * d);
  };
  block_sum_terms(term, nout, partial + (static_cast<size_t>(q) * gridDim.x + blockIdx.x) * nout);
}

// Sort keys of the queue, f n + g, for a deterministic order.
__global__ void dkey_kernel(const uint2 *__restrict__ queue, unsigned count, int n,
                            unsigned long long *__restrict__ key, unsigned *__restrict__ idx) {
  const unsigned e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= count) return;
  const uint2 qe = queue[e];
  key[e] = static_cast<unsigned long long>(qe.x & 0x3fffffffu) * static_cast<unsigned long long>(n) + qe.y;
  idx[e] = e;
}

// One wave per (sorted queue entry e, direction q): dl[q count + e], for the
// entries not keyed by aligned numbers (sampled pairs, heavy point
// quadratures of pairs that are not aligned).
__global__ void __launch_bounds__(64) ddeferred_kernel(const fh_filament *__restrict__ F, int n,
                                                       const fh_filament *__restrict__ T,
                                                       const uint8_t *__restrict__ moving,
                                                       const uint2 *__restrict__ queue,
                                                       const unsigned *__restrict__ sidx, unsigned count, int P,
                                                       const unsigned *__restrict__ hsorted,
                                                       const unsigned *__restrict__ unkeyed_begin,
                                                       double *__restrict__ dl, int *__restrict__ bad, Frame *__restrict__ ovf) {
  __shared__ Dual lds[kCloudCapacity];
  __shared__ DeferredScratch scratch_lds;
  const int lane = static_cast<int>(threadIdx.x);
  // the entries without aligned numbers close the hash order
  const unsigned beg = *unkeyed_begin;
  const size_t nun = count - beg;
  const size_t total = nun * P;
  for (size_t w = blockIdx.x; w < total; w += gridDim.x) {
    const unsigned e = hsorted[beg + static_cast<unsigned>(w % nun)];
    const int q = static_cast<int>(w / nun);
    const uint2 qe = queue[sidx[e]];
    const int f = static_cast<int>(qe.x & 0x3fffffffu), g = static_cast<int>(qe.y);
    const uint32_t kind = qe.x >> 30;
    const size_t qn = static_cast<size_t>(q) * n;
    double d = 0.0;
    if (moving[qn + f] || moving[qn + g]) {
      Plan p;
      plan_pair(F[f], T[qn + f], F[g], T[qn + g], p);
      bool ok;
      d = deferred_eval(p, kind, lds, scratch_lds, ovf + blockIdx.x * kStackOvf, lane, ok).d;
      if (lane == 0 && !ok) atomicOr(bad, 1);
      if (lane == 0 && p.unsupported) atomicOr(bad, 2);
    }
    if (lane == 0) dl[static_cast<size_t>(q) * count + e] = d;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Aligned pairs through their nine numbers.  An aligned pair's integral is a
// function I(a) of its AlignedBars numbers a alone (lengths, widths,
// heights, offsets; inductance/mod.rs:280-300), and with every branch
// decided on the values each dual operation is linear in the tangents, so
// the derivative along a direction q is sum_i dI/da_i da_i/dq.  I is even in
// each offset (a reflection of both bars through a plane of bar a) and
// symmetric in the two bars (with the offset negated), so the partials are
// formed once per canonical a -- offsets by magnitude, the two bars in
// (length, width, height) order -- and every (pair, direction) maps them
// back and takes only its own tangent da/dq from the pair plan: nine
// evaluations per distinct canonical a instead of one per pair and
// direction (graded cross-sections repeat the same relative positions; on
// the 6 156-filament spiral 1.63 M aligned entries have 0.49 M canonical
// numbers).

constexpr unsigned long long kNoAKey = ~0ull;

__device__ __forceinline__ void aligned_values(const Aligned &ab, double (&w)[9]) {
  w[0] = ab.length[0].v;
  w[1] = ab.length[1].v;
  w[2] = ab.width[0].v;
  w[3] = ab.width[1].v;
  w[4] = ab.height[0].v;
  w[5] = ab.height[1].v;
  w[6] = ab.offset[0].v;
  w[7] = ab.offset[1].v;
  w[8] = ab.offset[2].v;
}

// Canonical numbers of w and the map back: bit 0 set when the bars are
// swapped, bits 1-3 when offset 0-2 is negative.
__device__ __forceinline__ int canonical_aligned(const double (&w)[9], double (&c)[9]) {
  const bool swap = w[1] < w[0] || (w[1] == w[0] && (w[3] < w[2] || (w[3] == w[2] && w[5] < w[4])));
  c[0] = swap ? w[1] : w[0];
  c[1] = swap ? w[0] : w[1];
  c[2] = swap ? w[3] : w[2];
  c[3] = swap ? w[2] : w[3];
  c[4] = swap ? w[5] : w[4];
  c[5] = swap ? w[4] : w[5];
  int flags = swap ? 1 : 0;
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    c[6 + t] = __builtin_fabs(w[6 + t]);
    if (__builtin_signbit(w[6 + t])) flags |= 2 << t;
  }
  return flags;
}

// Sorted queue position e: for an aligned entry, or a heavy point
// quadrature of an aligned pair, the canonical aligned numbers, their hash
// with the kind (kNoAKey for every other entry) and the map back.
__global__ void akey_kernel(const fh_filament *__restrict__ F, const uint2 *__restrict__ queue,
                            const unsigned *__restrict__ sidx, unsigned count, double *__restrict__ keys,
                            unsigned long long *__restrict__ hash, unsigned *__restrict__ idx,
                            int8_t *__restrict__ aflags, int8_t *__restrict__ akind) {
  const unsigned e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= count) return;
  idx[e] = e;
  const uint2 qe = queue[sidx[e]];
  const uint32_t kind = qe.x >> 30;
  unsigned long long h = kNoAKey;
  int8_t fl = 0;
  Plan p;
  bool keyed = false;
  if (kind != kSampled) {
    const fh_filament zero = {};
    plan_pair(F[qe.x & 0x3fffffffu], zero, F[qe.y], zero, p);
    keyed = p.aligned;
  }
  if (keyed) {
    double w[9], c[9];
    aligned_values(p.ab, w);
    fl = static_cast<int8_t>(canonical_aligned(w, c));
    h = 0x9E3779B97F4A7C15ull ^ kind;
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      keys[static_cast<size_t>(e) * 9 + t] = c[t];
      h ^= static_cast<unsigned long long>(__double_as_longlong(c[t]));
      h ^= h >> 30;
      h *= 0xbf58476d1ce4e5b9ull;
      h ^= h >> 27;
      h *= 0x94d049bb133111ebull;
      h ^= h >> 31;
    }
    // the kind in the top two bits: the runs of each kind are contiguous in
    // hash order, light point quadratures first (kNoAKey sorts last).  A
    // light point quadrature's point count (at most kInlineEvals) comes
    // next, so that the threads of a wave in agrad_light_kernel run loops
    // of similar lengths.
    if (kind == kLightPoint) {
      const unsigned long long pts =
          static_cast<unsigned long long>(p.oa[0] * p.oa[1] * p.oa[2] * (p.ob[0] * p.ob[1] * p.ob[2]));
      h = (pts << 51) | (h >> 13);
    } else {
      h >>= 2;
    }
    h |= static_cast<unsigned long long>(kind) << 62;
  }
  hash[e] = h;
  aflags[e] = fl;
  akind[e] = static_cast<int8_t>(kind);
}

// head[j] = 1 where hash-sorted entry j starts a run of bit-equal keys.
__global__ void ahead_kernel(const unsigned long long *__restrict__ sh, const unsigned *__restrict__ si,
                             const double *__restrict__ keys, const int8_t *__restrict__ akind, unsigned count,
                             unsigned *__restrict__ head) {
  const unsigned j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  unsigned hd = 0;
  if (sh[j] != kNoAKey) {
    hd = 1;
    if (j > 0 && sh[j - 1] == sh[j]) {
      bool same = akind[si[j]] == akind[si[j - 1]];
      for (int t = 0; t < 9; ++t)
        same = same && __double_as_longlong(keys[static_cast<size_t>(si[j]) * 9 + t]) ==
                           __double_as_longlong(keys[static_cast<size_t>(si[j - 1]) * 9 + t]);
      if (same) hd = 0;
    }
  }
  head[j] = hd;
}

// urun[e] = the run (distinct key) of sorted queue position e, -1 for
// entries of the other kinds; uq[u] = the position whose key run u
// evaluates; *ucount = the number of runs.
__global__ void arun_kernel(const unsigned long long *__restrict__ sh, const unsigned *__restrict__ si,
                            const unsigned *__restrict__ head, const unsigned *__restrict__ run, unsigned count,
                            int *__restrict__ urun, unsigned *__restrict__ uq, unsigned *__restrict__ ucount,
                            unsigned *__restrict__ nlight, unsigned *__restrict__ unkeyed_begin) {
  const unsigned j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  const unsigned e = si[j];
  urun[e] = sh[j] == kNoAKey ? -1 : static_cast<int>(run[j]) - 1;
  if (head[j]) uq[run[j] - 1] = e;
  if (j == count - 1) *ucount = run[j];
  // the last light entry in hash order closes the light runs
  const bool light = sh[j] >> 62 == kLightPoint;
  if (light && (j == count - 1 || sh[j + 1] >> 62 != kLightPoint)) *nlight = run[j];
  if (sh[j] == kNoAKey && (j == 0 || sh[j - 1] != kNoAKey)) *unkeyed_begin = j;
}

// Bar of the canonical frame: centre (x, y, z), axes along x, y, z, full
// extents (l, w, h); tangent 1 on the one number `i` of nine: the centre's
// three at ix, ix + 1, ix + 2 (ix < 0: a fixed centre), the extents at il,
// iw, ih.
__device__ __forceinline__ Bar canonical_bar(double x, double y, double z, double l, double w, double h,
                                             int ix, int il, int iw, int ih, int i) {
  Bar b;
  const bool moves = ix >= 0;
  b.centre = V3{Dual{x, moves && i == ix ? 1.0 : 0.0}, Dual{y, moves && i == ix + 1 ? 1.0 : 0.0},
                Dual{z, moves && i == ix + 2 ? 1.0 : 0.0}};
  b.axes[0] = V3{cst(1.0), cst(0.0), cst(0.0)};
  b.axes[1] = V3{cst(0.0), cst(1.0), cst(0.0)};
  b.axes[2] = V3{cst(0.0), cst(0.0), cst(1.0)};
  b.half[0] = Dual{0.5 * l, i == il ? 0.5 : 0.0};
  b.half[1] = Dual{0.5 * w, i == iw ? 0.5 : 0.0};
  b.half[2] = Dual{0.5 * h, i == ih ? 0.5 : 0.0};
  return b;
}

// The point quadrature of run u's pair in the canonical frame (bar a at the
// origin, bar b at the offsets, both along x, y, z) with tangent 1 on
// canonical number i: the orders are the run's own pair's, b's own width and
// height orders moved to the canonical y/z axes it is aligned with, the bars
// trading places when the canonical numbers swap them.
__device__ __forceinline__ void canonical_point_plan(const fh_filament *__restrict__ F, uint2 qe, int flags,
                                                     const double *c, int i, Plan &p) {
  const fh_filament zero = {};
  Plan r;
  plan_pair(F[qe.x & 0x3fffffffu], zero, F[qe.y], zero, r);
  const bool sameb = __builtin_fabs(dot(r.a.axes[1], r.b.axes[1]).v) > 1.0 - 1e-12;
  const int obc0 = r.ob[0], obc1 = sameb ? r.ob[1] : r.ob[2], obc2 = sameb ? r.ob[2] : r.ob[1];
  const bool sw = flags & 1;
  p.oa[0] = sw ? obc0 : r.oa[0];
  p.oa[1] = sw ? obc1 : r.oa[1];
  p.oa[2] = sw ? obc2 : r.oa[2];
  p.ob[0] = sw ? r.oa[0] : obc0;
  p.ob[1] = sw ? r.oa[1] : obc1;
  p.ob[2] = sw ? r.oa[2] : obc2;
  p.a = canonical_bar(0.0, 0.0, 0.0, c[0], c[2], c[4], -1, 0, 2, 4, i);
  p.b = canonical_bar(c[6], c[7], c[8], c[1], c[3], c[5], 6, 1, 3, 5, i);
  p.cos = cst(1.0);
  p.point = true;
}

// One wave per (run u, canonical number i), u in [first, *ucount): the
// partial derivative in c_i (ugrad[9 u + i]) and the value (uval[u]) of the
// run's integral -- for an aligned entry I(c) of the aligned dispatch
// (scaled by the pair later); for a heavy point quadrature 1e-7 x the point
// sum of the canonical plan.
__global__ void __launch_bounds__(64) agrad_kernel(const fh_filament *__restrict__ F,
                                                   const uint2 *__restrict__ queue,
                                                   const unsigned *__restrict__ sidx,
                                                   const double *__restrict__ keys,
                                                   const int8_t *__restrict__ aflags,
                                                   const unsigned *__restrict__ uq,
                                                   const unsigned *__restrict__ ucount,
                                                   const unsigned *__restrict__ first,
                                                   double *__restrict__ uval, double *__restrict__ ugrad, Frame *__restrict__ ovf) {
  __shared__ Dual lds[kCloudCapacity];
  __shared__ DeferredScratch scratch_lds;
  const int lane = static_cast<int>(threadIdx.x);
  const size_t begin = static_cast<size_t>(*first) * 9, total = static_cast<size_t>(*ucount) * 9;
  for (size_t w = begin + blockIdx.x; w < total; w += gridDim.x) {
    const unsigned u = static_cast<unsigned>(w / 9);
    const int i = static_cast<int>(w % 9);
    const unsigned pos = uq[u];
    const double *c = keys + static_cast<size_t>(pos) * 9;
    const uint2 qe = queue[sidx[pos]];
    Dual v;
    if ((qe.x >> 30) == kAlignedKind) {
      Aligned ab;
      ab.length[0] = Dual{c[0], i == 0 ? 1.0 : 0.0};
      ab.length[1] = Dual{c[1], i == 1 ? 1.0 : 0.0};
      ab.width[0] = Dual{c[2], i == 2 ? 1.0 : 0.0};
      ab.width[1] = Dual{c[3], i == 3 ? 1.0 : 0.0};
      ab.height[0] = Dual{c[4], i == 4 ? 1.0 : 0.0};
      ab.height[1] = Dual{c[5], i == 5 ? 1.0 : 0.0};
      ab.offset[0] = Dual{c[6], i == 6 ? 1.0 : 0.0};
      ab.offset[1] = Dual{c[7], i == 7 ? 1.0 : 0.0};
      ab.offset[2] = Dual{c[8], i == 8 ? 1.0 : 0.0};
      v = aligned_dispatch_wave(ab, FrameStack{scratch_lds.stack, ovf + blockIdx.x * kStackOvf}, lds);
    } else {
      Plan p;
      canonical_point_plan(F, qe, aflags[pos], c, i, p);
      bool ok;
      v = deferred_eval(p, kHeavyPoint, lds, scratch_lds, ovf + blockIdx.x * kStackOvf, lane, ok);
    }
    if (lane == 0) {
      ugrad[w] = v.d;
      if (i == 0) uval[u] = v.v;
    }
    __syncthreads();
  }
}

// One thread per (run u, canonical number i), u < *nlight (the light point
// quadratures): 1e-7 x the serial point sum of the canonical plan, as the
// tile kernels evaluate a light pair.
// Three waves per SIMD (168 VGPRs, 42 spilled) rather than the two its 212
// registers allow: about 1 % faster on the spiral gradient, same bits.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) agrad_light_kernel(const fh_filament *__restrict__ F,
                                                          const uint2 *__restrict__ queue,
                                                          const unsigned *__restrict__ sidx,
                                                          const double *__restrict__ keys,
                                                          const int8_t *__restrict__ aflags,
                                                          const unsigned *__restrict__ uq,
                                                          const unsigned *__restrict__ nlight,
                                                          double *__restrict__ uval, double *__restrict__ ugrad) {
  const size_t total = static_cast<size_t>(*nlight) * 9;
  for (size_t w = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; w < total;
       w += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const unsigned u = static_cast<unsigned>(w / 9);
    const int i = static_cast<int>(w % 9);
    const unsigned pos = uq[u];
    Plan p;
    canonical_point_plan(F, queue[sidx[pos]], aflags[pos], keys + static_cast<size_t>(pos) * 9, i, p);
    const Dual v = 1.0e-7 * point_sum_serial(p.a, p.oa, p.b, p.ob);
    ugrad[w] = v.d;
    if (i == 0) uval[u] = v.v;
  }
}

// One thread per (sorted keyed entry e, direction q): dl[q count + e] =
// 1e-7 d(I sign(cos) / (area_a area_b)) for an aligned entry, d(I cos) for
// a heavy point quadrature, with dI = sum_i dI/da_i da_i/dq (the canonical
// partials mapped back through aflags[e]), as
// the fill scales an aligned integral (deferred_eval).
__global__ void __launch_bounds__(256) acontract_kernel(const fh_filament *__restrict__ F, int n,
                                                        const fh_filament *__restrict__ T,
                                                        const uint8_t *__restrict__ moving,
                                                        const uint2 *__restrict__ queue,
                                                        const unsigned *__restrict__ sidx, unsigned count, int P,
                                                        const int *__restrict__ urun,
                                                        const int8_t *__restrict__ aflags,
                                                        const double *__restrict__ uval,
                                                        const double *__restrict__ ugrad,
                                                        double *__restrict__ dl, int *__restrict__ bad) {
  const size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<size_t>(count) * P) return;
  const unsigned e = static_cast<unsigned>(t % count);
  const int q = static_cast<int>(t / count);
  const int u = urun[e];
  if (u < 0) return;
  const uint2 qe = queue[sidx[e]];
  const int f = static_cast<int>(qe.x & 0x3fffffffu), g = static_cast<int>(qe.y);
  const size_t qn = static_cast<size_t>(q) * n;
  double d = 0.0;
  if (moving[qn + f] || moving[qn + g]) {
    Plan p;
    plan_pair(F[f], T[qn + f], F[g], T[qn + g], p);
    if (p.unsupported) atomicOr(bad, 2);
    const double *gr = ugrad + static_cast<size_t>(u) * 9;
    const Aligned &ab = p.ab;
    // the canonical partials back on this pair's numbers
    const int fl = aflags[e];
    const bool sw = fl & 1;
    const double s0 = fl & 2 ? -1.0 : 1.0, s1 = fl & 4 ? -1.0 : 1.0, s2 = fl & 8 ? -1.0 : 1.0;
    const double di = (sw ? gr[1] : gr[0]) * ab.length[0].d + (sw ? gr[0] : gr[1]) * ab.length[1].d +
                      (sw ? gr[3] : gr[2]) * ab.width[0].d + (sw ? gr[2] : gr[3]) * ab.width[1].d +
                      (sw ? gr[5] : gr[4]) * ab.height[0].d + (sw ? gr[4] : gr[5]) * ab.height[1].d +
                      s0 * gr[6] * ab.offset[0].d + s1 * gr[7] * ab.offset[1].d + s2 * gr[8] * ab.offset[2].d;
    const Dual in{uval[u], di};
    if ((qe.x >> 30) == kAlignedKind) {
      const Dual scale = signum(p.cos.v) / (p.area_a * p.area_b);
      d = (1.0e-7 * (in * scale)).d;
    } else {
      d = (in * p.cos).d;  // point sums: 1e-7 already in
    }
  }
  dl[t] = d;
}

// Sorted entries in blocks of 256 per direction (blockIdx.y): weighted
// derivatives summed per block in a fixed order.
__global__ void __launch_bounds__(256) dcontract_kernel(const uint2 *__restrict__ queue,
                                                        const unsigned *__restrict__ sidx, unsigned count,
                                                        const double *__restrict__ dl, Weights W, int nout,
                                                        double2 *__restrict__ partial) {
  const unsigned e = blockIdx.x * 256 + threadIdx.x;
  const int q = blockIdx.y;
  double d = 0.0;
  int f = 0, g = 0;
  if (e < count) {
    d = dl[static_cast<size_t>(q) * count + e];
    if (d != 0.0) {
      const uint2 qe = queue[sidx[e]];
      f = static_cast<int>(qe.x & 0x3fffffffu);
      g = static_cast<int>(qe.y);
    }
  }
  auto term = [&](int o) {
    if (d == 0.0) return make_double2(0.0, 0.0);
    const double2 w = weight(W, f, g, o);
    return make_double2(w.x * d, w.y * d);
  };
  block_sum_terms(term, nout, partial + (static_cast<size_t>(q) * gridDim.x + blockIdx.x) * nout);
}

// out[q nout + o] = sum over blocks of partial[q][block][o] (fixed tree).
__global__ void __launch_bounds__(256) dreduce_kernel(const double2 *__restrict__ partial, int blocks, int nout,
                                                      double2 *__restrict__ out) {
  const int q = blockIdx.x / nout, o = blockIdx.x % nout;
  double re = 0.0, im = 0.0;
  for (int b = threadIdx.x; b < blocks; b += 256) {
    const double2 v = partial[(static_cast<size_t>(q) * blocks + b) * nout + o];
    re += v.x;
    im += v.y;
  }
  __shared__ double sre[256], sim[256];
  sre[threadIdx.x] = re;
  sim[threadIdx.x] = im;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s) {
      sre[threadIdx.x] += sre[threadIdx.x + s];
      sim[threadIdx.x] += sim[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = make_double2(sre[0], sim[0]);
}

// Single pairs (f, g, q): value and derivative, every method on one wave.
__global__ void __launch_bounds__(64) dpair_kernel(const fh_filament *__restrict__ F, int n,
                                                   const fh_filament *__restrict__ T,
                                                   const uint4 *__restrict__ pairs, size_t count,
                                                   double *__restrict__ value, double *__restrict__ deriv, Frame *__restrict__ ovf) {
  __shared__ Dual lds[kCloudCapacity];
  __shared__ DeferredScratch scratch_lds;
  const int lane = static_cast<int>(threadIdx.x);
  for (size_t e = blockIdx.x; e < count; e += gridDim.x) {
    const uint4 pr = pairs[e];
    const int f = static_cast<int>(pr.x), g = static_cast<int>(pr.y), q = static_cast<int>(pr.z);
    const size_t qn = static_cast<size_t>(q) * n;
    Plan p;
    plan_pair(F[f], T[qn + f], F[g], T[qn + g], p);
    Dual v = cst(0.0);
    bool ok = true;
    if (!p.orthogonal) {
      if (p.point && p.oa[0] * p.oa[1] * p.oa[2] * (p.ob[0] * p.ob[1] * p.ob[2]) <= kInlineEvals) {
        v = 1.0e-7 * (point_sum_serial(p.a, p.oa, p.b, p.ob) * p.cos);
      } else {
        const uint32_t kind = p.point ? kHeavyPoint : (p.aligned ? kAlignedKind : kSampled);
        v = deferred_eval(p, kind, lds, scratch_lds, ovf + blockIdx.x * kStackOvf, lane, ok);
      }
    }
    if (lane == 0) {
      value[e] = ok ? v.v : __builtin_nan("");
      deriv[e] = ok && !p.unsupported ? v.d : __builtin_nan("");
    }
    __syncthreads();
  }
}

// y[v][i] = sum_j L[i][j] x[v][j] for V complex vectors (real L, row-major):
// one workgroup per row, vectors four at a time, fixed-order reductions.
__global__ void __launch_bounds__(256) lmatvec_kernel(const double *__restrict__ L, int n,
                                                      const double2 *__restrict__ x, int V,
                                                      double2 *__restrict__ y) {
  const int i = blockIdx.x;
  const double *row = L + static_cast<size_t>(i) * n;
  __shared__ double2 red[4][4];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int v0 = 0; v0 < V; v0 += 4) {
    double2 acc[4] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0),
                      make_double2(0.0, 0.0)};
    for (int j = threadIdx.x; j < n; j += 256) {
      const double l = row[j];
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (v0 + c < V) {
          const double2 xv = x[static_cast<size_t>(v0 + c) * n + j];
          acc[c].x += l * xv.x;
          acc[c].y += l * xv.y;
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double re = acc[c].x, im = acc[c].y;
      for (int d = 32; d > 0; d >>= 1) {
        re += __shfl_xor(re, d);
        im += __shfl_xor(im, d);
      }
      if (lane == 0) red[wave][c] = make_double2(re, im);
    }
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < 4 && v0 + static_cast<int>(threadIdx.x) < V) {
      const int c = threadIdx.x;
      y[static_cast<size_t>(v0 + c) * n + i] =
          make_double2((red[0][c].x + red[1][c].x) + (red[2][c].x + red[3][c].x),
                       (red[0][c].y + red[1][c].y) + (red[2][c].y + red[3][c].y));
    }
    __syncthreads();
  }
}

fh_status upload_adjoint_tables() {
  static std::mutex mu;
  static std::vector<int> done;  // per device
  int dev = 0;
  FH_HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  if (static_cast<int>(done.size()) > dev && done[dev]) return FH_OK;
  double gx[kMaxRule + 1][kMaxRule] = {}, gw[kMaxRule + 1][kMaxRule] = {};
  for (int k = 1; k <= kMaxRule; ++k) gauss_rule_host(k, gx[k], gw[k]);
  // gauss.rs:488-495 order thresholds, as the fill's tables
  double pt[kPointMaxOrder], st[kSampleMaxOrder];
  const double lp = std::log(1.0 / 1e-8), ls = std::log(1.0 / 1e-6);
  for (int k = 1; k <= kPointMaxOrder; ++k) pt[k - 1] = std::sinh(lp / (2.0 * k));
  for (int k = 1; k <= kSampleMaxOrder; ++k) st[k - 1] = std::sinh(ls / (2.0 * k));
  FH_HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(a_gx), gx, sizeof(gx)));
  FH_HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(a_gw), gw, sizeof(gw)));
  FH_HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(a_point_th), pt, sizeof(pt)));
  FH_HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(a_sample_th), st, sizeof(st)));
  if (static_cast<int>(done.size()) <= dev) done.resize(dev + 1, 0);
  done[dev] = 1;
  return FH_OK;
}

unsigned wave_blocks() {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  return static_cast<unsigned>(cus) * 16;
}

}  // namespace

fh_status pair_derivatives(const fh_filament *d_F, int n, const fh_filament *d_T, const uint4 *d_pairs,
                           size_t count, double *d_value, double *d_deriv, hipStream_t stream) {
  if (count == 0) return FH_OK;
  fh_status st = upload_adjoint_tables();
  if (st) return st;
  (void)n;
  const unsigned blocks = static_cast<unsigned>(std::min<size_t>(count, wave_blocks()));
  DeviceArray<Frame> ovf;  // the stack frames beyond LDS, per wave
  if ((st = ovf.allocate(static_cast<size_t>(blocks) * kStackOvf))) return st;
  dpair_kernel<<<blocks, 64, 0, stream>>>(d_F, n, d_T, d_pairs, count, d_value, d_deriv, ovf.ptr);
  FH_HIP_TRY(hipGetLastError());
  FH_HIP_TRY(hipStreamSynchronize(stream));
  return FH_OK;
}

fh_status contract_inductance_derivative(const fh_filament *d_F, int n, const fh_filament *d_T, int P,
                                         const double2 *d_I, int K, int p, std::vector<double2> &out,
                                         hipStream_t stream) {
  const int nout = K * p * p;
  out.assign(static_cast<size_t>(P) * nout, make_double2(0.0, 0.0));
  if (n == 0 || P == 0 || nout == 0) return FH_OK;
  if (nout > kMaxOut) return set_error(FH_E_INVALID_ARGUMENT, "more than 128 frequency x port-pair outputs per pass");
  if (n >= (1 << 30)) return set_error(FH_E_INVALID_ARGUMENT, "too many filaments");
  fh_status st = upload_adjoint_tables();
  if (st) return st;
  const int nt = (n + kTile - 1) / kTile;
  const long long tiles = static_cast<long long>(nt) * (nt + 1) / 2;
  if (tiles > 0x7fffffffLL) return set_error(FH_E_INVALID_ARGUMENT, "too many filaments");
  const size_t pairs = static_cast<size_t>(n) * (n + 1) / 2;
  const Weights W{d_I, n, K, p};
  DeviceArray<uint8_t> moving;
  DeviceArray<double2> tpart, res_t;
  DeviceArray<uint2> queue;
  DeviceArray<unsigned> qcount;
  DeviceArray<int> flags;  // bit 0: divergent sampled pair, bit 1: unsupported turn
  DeviceArray<Frame> ovf;  // the aligned quadrature's stack frames beyond LDS, per wave
  const size_t nP = static_cast<size_t>(n) * P;
  if ((st = moving.allocate(nP)) || (st = tpart.allocate(static_cast<size_t>(P) * tiles * nout)) ||
      (st = queue.allocate(pairs)) || (st = qcount.allocate(1)) || (st = res_t.allocate(static_cast<size_t>(P) * nout)) ||
      (st = flags.allocate(1)))
    return st;
  FH_HIP_TRY(hipMemsetAsync(flags.ptr, 0, sizeof(int), stream));
  moving_kernel<<<static_cast<unsigned>((nP + 255) / 256), 256, 0, stream>>>(d_T, nP, moving.ptr);
  FH_HIP_TRY(hipGetLastError());
  FH_HIP_TRY(hipMemsetAsync(qcount.ptr, 0, sizeof(unsigned), stream));
  dtile_kernel<<<dim3(static_cast<unsigned>(tiles), static_cast<unsigned>(P)), 256, 0, stream>>>(
      d_F, n, nt, d_T, moving.ptr, W, nout, tpart.ptr, queue.ptr, qcount.ptr, flags.ptr);
  FH_HIP_TRY(hipGetLastError());
  dreduce_kernel<<<P * nout, 256, 0, stream>>>(tpart.ptr, static_cast<int>(tiles), nout, res_t.ptr);
  FH_HIP_TRY(hipGetLastError());
  unsigned count = 0;
  std::vector<double2> ht(static_cast<size_t>(P) * nout), hd(ht.size(), make_double2(0.0, 0.0));
  FH_HIP_TRY(hipMemcpyAsync(&count, qcount.ptr, sizeof(count), hipMemcpyDeviceToHost, stream));
  FH_HIP_TRY(hipMemcpyAsync(ht.data(), res_t.ptr, ht.size() * sizeof(double2), hipMemcpyDeviceToHost, stream));
  FH_HIP_TRY(hipStreamSynchronize(stream));
  if (count > 0) {
    DeviceArray<unsigned long long> key, skey;
    DeviceArray<unsigned> idx, sidx;
    DeviceArray<double> dl;
    DeviceArray<char> temp;
    DeviceArray<double2> cpart, res_d;
    if ((st = key.allocate(count)) || (st = skey.allocate(count)) || (st = idx.allocate(count)) ||
        (st = sidx.allocate(count)) || (st = dl.allocate(static_cast<size_t>(count) * P)))
      return st;
    dkey_kernel<<<(count + 255) / 256, 256, 0, stream>>>(queue.ptr, count, n, key.ptr, idx.ptr);
    FH_HIP_TRY(hipGetLastError());
    int end_bit = 1;
    while (end_bit < 64 && (static_cast<unsigned long long>(n) * n) >> end_bit) ++end_bit;
    size_t tb = 0;
    FH_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key.ptr, skey.ptr, idx.ptr, sidx.ptr,
                                                 static_cast<int>(count), 0, end_bit, stream));
    if ((st = temp.allocate(std::max<size_t>(tb, 1)))) return st;
    FH_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(temp.ptr, tb, key.ptr, skey.ptr, idx.ptr, sidx.ptr,
                                                 static_cast<int>(count), 0, end_bit, stream));
    const size_t waves = static_cast<size_t>(count) * P;
    // entries keyed by their canonical aligned numbers (aligned entries and
    // heavy point quadratures of aligned pairs): nine partials per distinct
    // key, then one dot product per (entry, direction); the other entries
    // one wave per (entry, direction)
    DeviceArray<double> akeys, uval, ugrad;
    DeviceArray<unsigned long long> ahash, sahash;
    DeviceArray<unsigned> aidx, saidx, head, run, uq, ucount, nlight, ubeg;
    DeviceArray<int> urun;
    DeviceArray<int8_t> aflags, akind;
    DeviceArray<char> atemp;
    if ((st = akeys.allocate(static_cast<size_t>(count) * 9)) || (st = ahash.allocate(count)) ||
        (st = sahash.allocate(count)) || (st = aidx.allocate(count)) || (st = saidx.allocate(count)) ||
        (st = head.allocate(count)) || (st = run.allocate(count)) || (st = uq.allocate(count)) ||
        (st = ucount.allocate(1)) || (st = urun.allocate(count)) || (st = uval.allocate(count)) ||
        (st = ugrad.allocate(static_cast<size_t>(count) * 9)) || (st = aflags.allocate(count)) ||
        (st = akind.allocate(count)) || (st = nlight.allocate(1)) || (st = ubeg.allocate(1)))
      return st;
    FH_HIP_TRY(hipMemsetAsync(nlight.ptr, 0, sizeof(unsigned), stream));
    // every entry keyed: the unkeyed range is empty
    FH_HIP_TRY(hipMemcpyAsync(ubeg.ptr, &count, sizeof(unsigned), hipMemcpyHostToDevice, stream));
    const unsigned eb = (count + 255) / 256;
    akey_kernel<<<eb, 256, 0, stream>>>(d_F, queue.ptr, sidx.ptr, count, akeys.ptr, ahash.ptr, aidx.ptr,
                                        aflags.ptr, akind.ptr);
    FH_HIP_TRY(hipGetLastError());
    size_t sort_b = 0, scan_b = 0;
    FH_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_b, ahash.ptr, sahash.ptr, aidx.ptr, saidx.ptr,
                                                 static_cast<int>(count), 0, 64, stream));
    FH_HIP_TRY(hipcub::DeviceScan::InclusiveSum(nullptr, scan_b, head.ptr, run.ptr, static_cast<int>(count),
                                                stream));
    size_t ab = std::max<size_t>(std::max(sort_b, scan_b), 1);
    if ((st = atemp.allocate(ab))) return st;
    FH_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(atemp.ptr, sort_b, ahash.ptr, sahash.ptr, aidx.ptr, saidx.ptr,
                                                 static_cast<int>(count), 0, 64, stream));
    ahead_kernel<<<eb, 256, 0, stream>>>(sahash.ptr, saidx.ptr, akeys.ptr, akind.ptr, count, head.ptr);
    FH_HIP_TRY(hipGetLastError());
    FH_HIP_TRY(hipcub::DeviceScan::InclusiveSum(atemp.ptr, ab, head.ptr, run.ptr, static_cast<int>(count),
                                                stream));
    arun_kernel<<<eb, 256, 0, stream>>>(sahash.ptr, saidx.ptr, head.ptr, run.ptr, count, urun.ptr, uq.ptr,
                                        ucount.ptr, nlight.ptr, ubeg.ptr);
    FH_HIP_TRY(hipGetLastError());
    const unsigned dblocks = static_cast<unsigned>(std::min<size_t>(waves, wave_blocks()));
    if ((st = ovf.allocate(static_cast<size_t>(wave_blocks()) * kStackOvf))) return st;
    ddeferred_kernel<<<dblocks, 64, 0, stream>>>(d_F, n, d_T, moving.ptr, queue.ptr, sidx.ptr, count, P,
                                                 saidx.ptr, ubeg.ptr, dl.ptr, flags.ptr, ovf.ptr);
    FH_HIP_TRY(hipGetLastError());
    agrad_kernel<<<wave_blocks(), 64, 0, stream>>>(d_F, queue.ptr, sidx.ptr, akeys.ptr, aflags.ptr, uq.ptr,
                                                   ucount.ptr, nlight.ptr, uval.ptr, ugrad.ptr, ovf.ptr);
    FH_HIP_TRY(hipGetLastError());
    agrad_light_kernel<<<wave_blocks(), 256, 0, stream>>>(d_F, queue.ptr, sidx.ptr, akeys.ptr, aflags.ptr,
                                                          uq.ptr, nlight.ptr, uval.ptr, ugrad.ptr);
    FH_HIP_TRY(hipGetLastError());
    acontract_kernel<<<static_cast<unsigned>((waves + 255) / 256), 256, 0, stream>>>(
        d_F, n, d_T, moving.ptr, queue.ptr, sidx.ptr, count, P, urun.ptr, aflags.ptr, uval.ptr, ugrad.ptr,
        dl.ptr, flags.ptr);
    FH_HIP_TRY(hipGetLastError());
    const unsigned cblocks = (count + 255) / 256;
    if ((st = cpart.allocate(static_cast<size_t>(P) * cblocks * nout)) ||
        (st = res_d.allocate(static_cast<size_t>(P) * nout)))
      return st;
    dcontract_kernel<<<dim3(cblocks, static_cast<unsigned>(P)), 256, 0, stream>>>(queue.ptr, sidx.ptr, count,
                                                                                 dl.ptr, W, nout, cpart.ptr);
    FH_HIP_TRY(hipGetLastError());
    dreduce_kernel<<<P * nout, 256, 0, stream>>>(cpart.ptr, static_cast<int>(cblocks), nout, res_d.ptr);
    FH_HIP_TRY(hipGetLastError());
    FH_HIP_TRY(hipMemcpyAsync(hd.data(), res_d.ptr, hd.size() * sizeof(double2), hipMemcpyDeviceToHost, stream));
    FH_HIP_TRY(hipStreamSynchronize(stream));  // the temporaries are freed at the end of this scope
  }
  int hflags = 0;
  FH_HIP_TRY(hipMemcpyAsync(&hflags, flags.ptr, sizeof(int), hipMemcpyDeviceToHost, stream));
  FH_HIP_TRY(hipStreamSynchronize(stream));
  if (hflags & 1) return set_error(FH_E_KERNEL, "a sampled-filament pair has coincident sample lines", 0, 0, 0.0, 1);
  if (hflags & 2)
    return set_error(FH_E_UNSUPPORTED,
                     "a direction turns a perpendicular or parallel filament pair relative to each other; "
                     "only directions that move and resize filaments without turning them are supported");
  for (size_t i = 0; i < out.size(); ++i) out[i] = make_double2(ht[i].x + hd[i].x, ht[i].y + hd[i].y);
  return FH_OK;
}

fh_status inductance_products(const double *d_L, int n, const double2 *d_I, int V, std::vector<double2> &out,
                              hipStream_t stream) {
  out.assign(static_cast<size_t>(V) * n, make_double2(0.0, 0.0));
  if (n == 0 || V == 0) return FH_OK;
  DeviceArray<double2> y;
  fh_status st;
  if ((st = y.allocate(static_cast<size_t>(V) * n))) return st;
  lmatvec_kernel<<<n, 256, 0, stream>>>(d_L, n, d_I, V, y.ptr);
  FH_HIP_TRY(hipGetLastError());
  FH_HIP_TRY(hipMemcpyAsync(out.data(), y.ptr, out.size() * sizeof(double2), hipMemcpyDeviceToHost, stream));
  FH_HIP_TRY(hipStreamSynchronize(stream));
  return FH_OK;
}

}  // namespace fh
