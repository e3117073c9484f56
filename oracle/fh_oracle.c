/*
 * fh_oracle.c -- CPU FP64 restatement of fasterhenry's extraction hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see fh_oracle.h): the checker for the GPU
 * product and the CPU baseline of bench.py.  It is written to follow the
 * reference's arithmetic operation by operation (same operand order, no
 * fused multiply-add: build with -ffp-contract=off), so that, with the
 * portable math mode, the GPU fill reproduces its bits.
 *
 * Reference paths are relative to the reference repository root.  Rust's
 * f64 methods map as follows: ln -> log, atan, hypot, sqrt; nalgebra's
 * Vector3 dot is (x*x' + y*y') + z*z', norm = sqrt(dot(v, v)); Rust's powi
 * is compiler-rt's binary powering; `wide::f64x4::reduce_add` on a default
 * x86-64 (SSE2) build is (l0 + l1) + (l2 + l3).
 */
#include "fh_oracle.h"

#include <math.h>
#include <omp.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Error state                                                               */
/* ------------------------------------------------------------------------ */

static __thread fho_error g_error;

static int fail(int status, int code, int64_t i0, int64_t i1, double value) {
  g_error.code = code;
  g_error.index0 = i0;
  g_error.index1 = i1;
  g_error.value = value;
  return status;
}

void fho_last_error(fho_error *out) { *out = g_error; }

/* ------------------------------------------------------------------------ */
/* Transcendental functions                                                  */
/* ------------------------------------------------------------------------ */

static int g_portable = 1;
void fho_set_math(int portable) { g_portable = portable ? 1 : 0; }
int fho_get_math(void) { return g_portable; }

static inline uint64_t bits_of(double x) {
  uint64_t b;
  memcpy(&b, &x, 8);
  return b;
}
static inline double double_of(uint64_t b) {
  double x;
  memcpy(&x, &b, 8);
  return x;
}

/* Natural logarithm from IEEE basic operations only.
 * x = 2^k m, m in [sqrt(1/2), sqrt(2)), f = m - 1 (exact), s = f/(2 + f),
 * log(1 + f) = 2 atanh(s) = f - f^2/2 + s (f^2/2 + R(s^2)) with
 * R(z) = sum_k 2 z^k / (2k + 1) (the atanh series; |s| <= 0.1716, eleven
 * terms leave a truncation error below 1e-18), and k ln 2 added as a
 * 32-bit head (exact product) plus tail. */
static double portable_log(double x) {
  if (!(x > 0.0)) {
    if (x == 0.0) return -INFINITY;
    return NAN; /* negative or NaN */
  }
  if (x == INFINITY) return x;
  uint64_t b = bits_of(x);
  int k = 0;
  if (b < 0x0010000000000000ULL) { /* subnormal */
    x *= 0x1p54;
    b = bits_of(x);
    k = -54;
  }
  k += (int)(b >> 52) - 1023;
  double m = double_of((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
  if (m > 0x1.6a09e667f3bcdp+0) {
    m *= 0.5;
    k += 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  const double r =
      z * (2.0 / 3.0 +
           z * (2.0 / 5.0 +
                z * (2.0 / 7.0 +
                     z * (2.0 / 9.0 +
                          z * (2.0 / 11.0 +
                               z * (2.0 / 13.0 +
                                    z * (2.0 / 15.0 +
                                         z * (2.0 / 17.0 +
                                              z * (2.0 / 19.0 +
                                                   z * (2.0 / 21.0 + z * (2.0 / 23.0)))))))))));
  const double hfsq = 0.5 * f * f;
  const double dk = (double)k;
  const double ln2_hi = 0x1.62e42feep-1, ln2_lo = 0x1.a39ef35793c76p-33;
  return dk * ln2_hi - ((hfsq - (s * (hfsq + r) + dk * ln2_lo)) - f);
}

/* Arctangent from IEEE basic operations only.  |x| is reduced about
 * c in {0, 1/2, 1, 3/2, inf} by atan(x) = atan(c) + atan((x - c)/(1 + c x)),
 * leaving |u| <= 7/16, where the Taylor series u - u^3/3 + u^5/5 - ... is
 * summed to 25 terms (truncation below 1e-19).  atan(c) is a head/tail pair
 * (values from 113-bit arithmetic). */
static double portable_atan(double x) {
  if (x != x) return x;
  const double ax = fabs(x);
  static const double hi[4] = {0x1.dac670561bb4fp-2, 0x1.921fb54442d18p-1, 0x1.f730bd281f69bp-1,
                               0x1.921fb54442d18p+0};
  static const double lo[4] = {0x1.a2b7f222f65e2p-56, 0x1.1a62633145c07p-55,
                               0x1.007887af0cbbdp-56, 0x1.1a62633145c07p-54};
  if (ax >= 0x1p66) {
    const double r = hi[3] + lo[3];
    return x < 0 ? -r : r;
  }
  int id;
  double u;
  if (ax < 0.4375) {
    if (ax < 0x1p-27) return x;
    id = -1;
    u = ax;
  } else if (ax < 0.6875) {
    id = 0;
    u = (2.0 * ax - 1.0) / (2.0 + ax);
  } else if (ax < 1.1875) {
    id = 1;
    u = (ax - 1.0) / (ax + 1.0);
  } else if (ax < 2.4375) {
    id = 2;
    u = (ax - 1.5) / (1.0 + 1.5 * ax);
  } else {
    id = 3;
    u = -1.0 / ax;
  }
  const double z = u * u;
  double p = 1.0 / 49.0;
  p = -1.0 / 47.0 + z * p;
  p = 1.0 / 45.0 + z * p;
  p = -1.0 / 43.0 + z * p;
  p = 1.0 / 41.0 + z * p;
  p = -1.0 / 39.0 + z * p;
  p = 1.0 / 37.0 + z * p;
  p = -1.0 / 35.0 + z * p;
  p = 1.0 / 33.0 + z * p;
  p = -1.0 / 31.0 + z * p;
  p = 1.0 / 29.0 + z * p;
  p = -1.0 / 27.0 + z * p;
  p = 1.0 / 25.0 + z * p;
  p = -1.0 / 23.0 + z * p;
  p = 1.0 / 21.0 + z * p;
  p = -1.0 / 19.0 + z * p;
  p = 1.0 / 17.0 + z * p;
  p = -1.0 / 15.0 + z * p;
  p = 1.0 / 13.0 + z * p;
  p = -1.0 / 11.0 + z * p;
  p = 1.0 / 9.0 + z * p;
  p = -1.0 / 7.0 + z * p;
  p = 1.0 / 5.0 + z * p;
  p = -1.0 / 3.0 + z * p;
  p = z * p; /* atan(u) = u + u * p */
  double r;
  if (id < 0) {
    r = u + u * p;
  } else {
    r = hi[id] - ((-(u * p) - lo[id]) - u);
  }
  return x < 0 ? -r : r;
}

double fho_log(double x) { return g_portable ? portable_log(x) : log(x); }
double fho_atan(double x) { return g_portable ? portable_atan(x) : atan(x); }
double fho_hypot(double x, double y) { return g_portable ? sqrt(x * x + y * y) : hypot(x, y); }

#define LN(x) fho_log(x)
#define ATAN(x) fho_atan(x)
#define HYPOT(x, y) fho_hypot(x, y)

/* ------------------------------------------------------------------------ */
/* Vector helpers with nalgebra's operation order                            */
/* ------------------------------------------------------------------------ */

typedef struct {
  double x, y, z;
} v3;

static inline v3 V(double x, double y, double z) {
  v3 r = {x, y, z};
  return r;
}
static inline v3 vload(const double *p) { return V(p[0], p[1], p[2]); }
static inline void vstore(double *p, v3 a) {
  p[0] = a.x;
  p[1] = a.y;
  p[2] = a.z;
}
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, double s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 vdiv(v3 a, double s) { return V(a.x / s, a.y / s, a.z / s); }
static inline double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b) {
  return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double vnorm(v3 a) { return sqrt(vdot(a, a)); }
static inline v3 vabs(v3 a) { return V(fabs(a.x), fabs(a.y), fabs(a.z)); }
/* Rust f64::max / min: a NaN argument yields the other one. */
static inline double fmax_r(double a, double b) { return fmax(a, b); }
static inline double fmin_r(double a, double b) { return fmin(a, b); }
static inline double signum(double x) {
  if (x != x) return x;
  return signbit(x) ? -1.0 : 1.0;
}
/* compiler-rt __powidf2, which Rust's powi lowers to for a variable exponent */
static double powi_r(double a, int b) {
  const int recip = b < 0;
  double r = 1.0;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1.0 / r : r;
}

/* ------------------------------------------------------------------------ */
/* Gauss-Legendre rules (gauss.rs:25-70) and order tables (gauss.rs:100-137) */
/* ------------------------------------------------------------------------ */

#define GAUSS_MAX 64
static double g_nodes[GAUSS_MAX + 1][GAUSS_MAX];
static double g_weights[GAUSS_MAX + 1][GAUSS_MAX];

static void legendre(int n, double x, double *p_out, double *p_prev_out) {
  double p_prev = 1.0, p = x;
  for (int j = 1; j < n; ++j) {
    const double jf = (double)j;
    const double p_next = ((2.0 * jf + 1.0) * x * p - jf * p_prev) / (jf + 1.0);
    p_prev = p;
    p = p_next;
  }
  *p_out = p;
  *p_prev_out = p_prev;
}

static void rule_new(int n, double *nodes, double *weights) {
  const double nf = (double)n;
  for (int k = 0; k < (n + 1) / 2; ++k) {
    double x = cos(3.141592653589793 * ((double)k + 0.75) / (nf + 0.5));
    double derivative = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p, p_prev;
      legendre(n, x, &p, &p_prev);
      derivative = nf * (x * p - p_prev) / (x * x - 1.0);
      const double step = p / derivative;
      x -= step;
      if (fabs(step) <= 4.0 * 2.220446049250313e-16 * fmax_r(fabs(x), 1e-3)) break;
    }
    if (n % 2 == 1 && k == n / 2) {
      x = 0.0;
      double p, p_prev;
      legendre(n, x, &p, &p_prev);
      derivative = nf * (x * p - p_prev) / (x * x - 1.0);
    }
    const double weight = 2.0 / ((1.0 - x * x) * derivative * derivative);
    nodes[n - 1 - k] = x;
    nodes[k] = -x;
    weights[n - 1 - k] = weight;
    weights[k] = weight;
  }
}

/* thresholds[n-1] = sinh(ln(1/tol) / 2n) (gauss.rs:488-495) */
#define POINT_MAX_ORDER 16
#define SAMPLE_MAX_ORDER 8
static double g_point_thresholds[POINT_MAX_ORDER];
static double g_sample_thresholds[SAMPLE_MAX_ORDER];

static void table_new(double tolerance, int max_order, double *thresholds) {
  const double l = log(1.0 / tolerance);
  for (int n = 1; n <= max_order; ++n) thresholds[n - 1] = sinh(l / (2.0 * (double)n));
}

static pthread_once_t g_once = PTHREAD_ONCE_INIT;
static void init_tables(void) {
  for (int n = 1; n <= GAUSS_MAX; ++n) rule_new(n, g_nodes[n], g_weights[n]);
  table_new(1e-8, POINT_MAX_ORDER, g_point_thresholds);
  table_new(1e-6, SAMPLE_MAX_ORDER, g_sample_thresholds);
}
static inline void ensure_tables(void) { pthread_once(&g_once, init_tables); }

void fho_gauss_rule(int n, double *nodes, double *weights) {
  ensure_tables();
  if (n < 1) n = 1;
  if (n > GAUSS_MAX) n = GAUSS_MAX;
  memcpy(nodes, g_nodes[n], sizeof(double) * n);
  memcpy(weights, g_weights[n], sizeof(double) * n);
}

/* order_for (gauss.rs:504-517); 0 means None. */
static int order_for(const double *thresholds, int max_order, double half_extent,
                     double distance) {
  if (half_extent <= 0.0) return 1;
  if (distance != distance || distance <= 0.0) return 0;
  const double delta = distance / half_extent;
  for (int i = 0; i < max_order; ++i)
    if (delta >= thresholds[i]) return i + 1;
  return 0;
}

int fho_order_for(double tolerance, int max_order, double half_extent, double distance) {
  double th[GAUSS_MAX];
  if (max_order < 1 || max_order > GAUSS_MAX) return 0;
  table_new(tolerance, max_order, th);
  return order_for(th, max_order, half_extent, distance);
}

/* Rule::on (gauss.rs:452-458): node and weight i mapped onto [a, b]. */
static inline void rule_on(int n, int i, double a, double b, double *x, double *w) {
  const double mid = 0.5 * (a + b), half = 0.5 * (b - a);
  *x = mid + half * g_nodes[n][i];
  *w = half * g_weights[n][i];
}

/* ------------------------------------------------------------------------ */
/* bar.rs: 64-term Hoer-Love closed form                                     */
/* ------------------------------------------------------------------------ */

void fho_corner_differences(double e1, double e2, double offset, double out[4]) {
  const double sum = 0.5 * (e1 + e2), diff = 0.5 * (e1 - e2);
  out[0] = offset + sum;
  out[1] = offset + diff;
  out[2] = offset - diff;
  out[3] = offset - sum;
}

static const double SIGNS[4] = {1.0, -1.0, -1.0, 1.0};

/* bar.rs:41-48 */
static double log_term(double coefficient, double x, double r, double rest_sq) {
  if (coefficient == 0.0 || x == 0.0) return 0.0;
  const double argument = x >= 0.0 ? x + r : rest_sq / (r - x);
  return coefficient * x * LN(argument);
}

/* bar.rs:51-57 */
static double atan_term(double x, double y, double z, double r) {
  const double product = x * y * z;
  if (product == 0.0) return 0.0;
  return -(product * z * z / 6.0) * ATAN(x * y / (z * r));
}

/* bar.rs:60-71 */
static double primitive(double x, double y, double z) {
  const double x2 = x * x, y2 = y * y, z2 = z * z;
  const double x4 = x2 * x2, y4 = y2 * y2, z4 = z2 * z2;
  const double r = sqrt(x2 + y2 + z2);
  return log_term(y2 * z2 / 4.0 - y4 / 24.0 - z4 / 24.0, x, r, y2 + z2) +
         log_term(x2 * z2 / 4.0 - x4 / 24.0 - z4 / 24.0, y, r, x2 + z2) +
         log_term(x2 * y2 / 4.0 - x4 / 24.0 - y4 / 24.0, z, r, x2 + y2) +
         (x4 + y4 + z4 - 3.0 * (x2 * y2 + y2 * z2 + z2 * x2)) * r / 60.0 +
         atan_term(x, y, z, r) + atan_term(x, z, y, r) + atan_term(y, z, x, r);
}

/* bar.rs:84-106 */
double fho_bar_integral(const double qx[4], const double qy[4], const double qz[4],
                        double *relative_error) {
  double value = 0.0, magnitude = 0.0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 4; ++k) {
        const double term = SIGNS[i] * SIGNS[j] * SIGNS[k] * primitive(qx[i], qy[j], qz[k]);
        value += term;
        magnitude += fabs(term);
      }
  if (relative_error)
    *relative_error = value > 0.0 ? 8.0 * 2.220446049250313e-16 * magnitude / value : INFINITY;
  return value;
}

/* ------------------------------------------------------------------------ */
/* lines.rs: Neumann line-integral closed forms                              */
/* ------------------------------------------------------------------------ */

/* lines.rs:31-40 */
static double parallel_primitive(double q, double rho_sq) {
  q = fabs(q);
  const double root = sqrt(q * q + rho_sq);
  if (q == 0.0) return -root;
  return q * LN(q + root) - root;
}

/* lines.rs:43-45 */
static double axial_overlap(double l1, double l2, double offset) {
  return fmax_r(fmin_r(l1, offset + l2) - fmax_r(offset, 0.0), 0.0);
}

/* lines.rs:54-70 */
int fho_lines_parallel(double l1, double l2, double offset, double rho, double *out) {
  const double rho_sq = rho * rho;
  const double smooth = parallel_primitive(offset + l2, rho_sq) -
                        parallel_primitive(offset, rho_sq) -
                        parallel_primitive(offset + l2 - l1, rho_sq) +
                        parallel_primitive(offset - l1, rho_sq);
  const double overlap = axial_overlap(l1, l2, offset);
  if (overlap == 0.0) {
    *out = smooth;
    return 1;
  }
  if (rho > 0.0) {
    *out = smooth - 2.0 * overlap * LN(rho);
    return 1;
  }
  return 0;
}

/* lines.rs:74-90 */
static double skew_log_term(double x, double big_r, double a, double residual_sq) {
  if (x == 0.0) return 0.0;
  const double argument = a >= 0.0 ? big_r + a : residual_sq / (big_r - a);
  if (argument > 0.0) return x * LN(argument);
  return 0.0;
}

/* lines.rs:98-124 */
static double skew_primitive(double s, double t, double d, double cos_, double sin_) {
  const double sin_sq = sin_ * sin_;
  const double versine = sin_sq / (1.0 + fabs(cos_));
  double planar_sq, a_s, a_t;
  if (cos_ >= 0.0) {
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
  const double big_r = sqrt(fmax_r(planar_sq + d * d, 0.0));
  const double logs = skew_log_term(s, big_r, a_s, s * s * sin_sq + d * d) +
                      skew_log_term(t, big_r, a_t, t * t * sin_sq + d * d);
  if (d == 0.0 || big_r == 0.0) return logs;
  return logs - (d / sin_) * ATAN((d * d * cos_ + s * t * sin_sq) / (d * big_r * sin_));
}

/* lines.rs:135-170 */
static int skew_v(v3 a0, v3 a1, v3 b0, v3 b1, double sin_min, double *out) {
  const double la = vnorm(vsub(a1, a0)), lb = vnorm(vsub(b1, b0));
  const v3 u = vdiv(vsub(a1, a0), la), v = vdiv(vsub(b1, b0), lb);
  const double cos_ = vdot(u, v);
  const double sin_ = vnorm(vcross(u, v));
  if (sin_ != sin_ || sin_ < sin_min) return 0;
  const v3 w0 = vsub(a0, b0);
  const double du = vdot(u, w0), dv = vdot(v, w0);
  const double sin_sq = sin_ * sin_;
  const double s0 = (cos_ * dv - du) / sin_sq;
  const double t0 = (dv - cos_ * du) / sin_sq;
  double d = vnorm(vsub(vadd(w0, vmul(u, s0)), vmul(v, t0)));
  if (d <= 1e-14 * (la + lb + vnorm(w0))) d = 0.0;
  const double s_lo = -s0, s_hi = la - s0;
  const double t_lo = -t0, t_hi = lb - t0;
  *out = skew_primitive(s_hi, t_hi, d, cos_, sin_) - skew_primitive(s_hi, t_lo, d, cos_, sin_) -
         skew_primitive(s_lo, t_hi, d, cos_, sin_) + skew_primitive(s_lo, t_lo, d, cos_, sin_);
  return 1;
}

int fho_lines_skew(const double a0[3], const double a1[3], const double b0[3],
                   const double b1[3], double sin_min, double *out) {
  return skew_v(vload(a0), vload(a1), vload(b0), vload(b1), sin_min, out);
}

/* ------------------------------------------------------------------------ */
/* aligned.rs: semi-analytic aligned-bar integral                            */
/* ------------------------------------------------------------------------ */

#define FAR_ORDER 10
#define CORNER_ORDER 16
#define SNAP 1e-12

typedef struct {
  double length[2], width[2], height[2], offset[3];
} aligned_bars;

/* aligned.rs:78-82 */
static double overlap_weight(double e1, double e2, double offset, double u) {
  const double hi = fmin_r(0.5 * e1, offset + 0.5 * e2 - u);
  const double lo = fmax_r(-0.5 * e1, offset - 0.5 * e2 - u);
  return fmax_r(hi - lo, 0.0);
}

static int cmp_total(const void *pa, const void *pb) {
  int64_t a = (int64_t)bits_of(*(const double *)pa), b = (int64_t)bits_of(*(const double *)pb);
  a ^= (int64_t)(((uint64_t)(a >> 63)) >> 1);
  b ^= (int64_t)(((uint64_t)(b >> 63)) >> 1);
  return (a > b) - (a < b);
}

/* aligned.rs:86-101; returns the count (<= 5). */
static int break_points(double e1, double e2, double offset, double *points) {
  const double sum = 0.5 * (e1 + e2), diff = 0.5 * fabs(e1 - e2);
#define SNAPV(x) (fabs(x) <= SNAP * sum ? 0.0 : (x))
  double p[5];
  int n = 4;
  p[0] = SNAPV(offset - sum);
  p[1] = SNAPV(offset - diff);
  p[2] = SNAPV(offset + diff);
  p[3] = SNAPV(offset + sum);
#undef SNAPV
  if (p[0] < 0.0 && p[3] > 0.0) p[n++] = 0.0;
  qsort(p, (size_t)n, sizeof(double), cmp_total);
  int m = 0;
  for (int i = 0; i < n; ++i)
    if (m == 0 || !(points[m - 1] == p[i])) points[m++] = p[i];
  return m;
}

typedef struct {
  const aligned_bars *bars;
  double axial_start, axial_scale;
} integrator;

/* aligned.rs:112-122 */
static double integrand(const integrator *it, double u, double v) {
  const aligned_bars *b = it->bars;
  const double weight = overlap_weight(b->width[0], b->width[1], b->offset[1], u) *
                        overlap_weight(b->height[0], b->height[1], b->offset[2], v);
  if (weight == 0.0) return 0.0;
  double kernel;
  if (!fho_lines_parallel(b->length[0], b->length[1], it->axial_start, HYPOT(u, v), &kernel))
    kernel = 0.0;
  return weight * kernel;
}

/* aligned.rs:125-134 */
static double direct(const integrator *it, const double u[2], const double v[2]) {
  ensure_tables();
  double sum = 0.0;
  for (int i = 0; i < FAR_ORDER; ++i) {
    double uu, wu;
    rule_on(FAR_ORDER, i, u[0], u[1], &uu, &wu);
    for (int j = 0; j < FAR_ORDER; ++j) {
      double vv, wv;
      rule_on(FAR_ORDER, j, v[0], v[1], &vv, &wv);
      sum += wu * wv * integrand(it, uu, vv);
    }
  }
  return sum;
}

/* aligned.rs:138-151 */
static double corner(const integrator *it, double a, double b) {
  double sum = 0.0;
  for (int i = 0; i < CORNER_ORDER; ++i) {
    double tau, w_tau;
    rule_on(CORNER_ORDER, i, 0.0, 1.0, &tau, &w_tau);
    const double s = tau * tau * tau;
    const double radial = w_tau * 3.0 * tau * tau * s;
    for (int j = 0; j < CORNER_ORDER; ++j) {
      double t, w_t;
      rule_on(CORNER_ORDER, j, 0.0, 1.0, &t, &w_t);
      const double value = integrand(it, a * s, b * s * t) + integrand(it, a * s * t, b * s);
      sum += radial * w_t * value;
    }
  }
  return sum * fabs(a * b);
}

static double gap_of(const double r[2]) {
  if (r[0] > 0.0) return r[0];
  if (r[1] < 0.0) return -r[1];
  return 0.0;
}

/* aligned.rs:153-199 */
static double panel(const integrator *it, const double u[2], const double v[2], unsigned depth) {
  const double su = u[1] - u[0], sv = v[1] - v[0];
  if (su <= 0.0 || sv <= 0.0) return 0.0;
  const double distance = HYPOT(gap_of(u), gap_of(v));
  const double size = fmax_r(su, sv);
  if (depth >= 200) return direct(it, u, v);
#define SPLIT_U(at)                                                                  \
  do {                                                                               \
    const double a_ = (at);                                                          \
    const double ul[2] = {u[0], a_}, ur[2] = {a_, u[1]};                             \
    return panel(it, ul, v, depth + 1) + panel(it, ur, v, depth + 1);                \
  } while (0)
#define SPLIT_V(at)                                                                  \
  do {                                                                               \
    const double a_ = (at);                                                          \
    const double vl[2] = {v[0], a_}, vr[2] = {a_, v[1]};                             \
    return panel(it, u, vl, depth + 1) + panel(it, u, vr, depth + 1);                \
  } while (0)
  if (distance > 0.0) {
    if (size <= distance) return direct(it, u, v);
    if (su >= sv) SPLIT_U(0.5 * (u[0] + u[1]));
    SPLIT_V(0.5 * (v[0] + v[1]));
  }
  const double target = fmin_r(fmin_r(su, sv), it->axial_scale);
  if (su > 2.0 * target) SPLIT_U(u[0] == 0.0 ? target : -target);
  if (sv > 2.0 * target) SPLIT_V(v[0] == 0.0 ? target : -target);
#undef SPLIT_U
#undef SPLIT_V
  return corner(it, u[0] == 0.0 ? u[1] : u[0], v[0] == 0.0 ? v[1] : v[0]);
}

/* aligned.rs:203-232 */
static double aligned_integral_s(const aligned_bars *bars) {
  ensure_tables();
  const double axial_start = bars->offset[0] + 0.5 * (bars->length[0] - bars->length[1]);
  const double l1 = bars->length[0], l2 = bars->length[1];
  const double separations[4] = {axial_start + l2, axial_start, axial_start + l2 - l1,
                                 axial_start - l1};
  double largest = 0.0;
  for (int i = 0; i < 4; ++i) largest = fmax_r(largest, fabs(separations[i]));
  double axial_scale = INFINITY;
  for (int i = 0; i < 4; ++i) {
    const double q = fabs(separations[i]);
    if (q > 1e-9 * largest) axial_scale = fmin_r(axial_scale, q);
  }
  integrator it = {bars, axial_start, axial_scale};
  double us[5], vs[5];
  const int nu = break_points(bars->width[0], bars->width[1], bars->offset[1], us);
  const int nv = break_points(bars->height[0], bars->height[1], bars->offset[2], vs);
  double sum = 0.0;
  for (int i = 0; i + 1 < nu; ++i)
    for (int j = 0; j + 1 < nv; ++j) {
      const double u[2] = {us[i], us[i + 1]}, v[2] = {vs[j], vs[j + 1]};
      sum += panel(&it, u, v, 0);
    }
  return sum;
}

static aligned_bars bars_of(const double b[9]) {
  aligned_bars r = {{b[0], b[1]}, {b[2], b[3]}, {b[4], b[5]}, {b[6], b[7], b[8]}};
  return r;
}

double fho_aligned_integral(const double bars[9]) {
  const aligned_bars b = bars_of(bars);
  return aligned_integral_s(&b);
}

/* mod.rs:265-276 */
static double aligned_dispatch_s(const aligned_bars *b, int *method) {
  double qx[4], qy[4], qz[4], err;
  fho_corner_differences(b->length[0], b->length[1], b->offset[0], qx);
  fho_corner_differences(b->width[0], b->width[1], b->offset[1], qy);
  fho_corner_differences(b->height[0], b->height[1], b->offset[2], qz);
  const double exact = fho_bar_integral(qx, qy, qz, &err);
  if (err <= 1e-10) {
    *method = FHO_METHOD_BAR_CLOSED_FORM;
    return exact;
  }
  *method = FHO_METHOD_ALIGNED_QUADRATURE;
  return aligned_integral_s(b);
}

double fho_aligned_dispatch(const double bars[9], int *method) {
  const aligned_bars b = bars_of(bars);
  int m;
  const double r = aligned_dispatch_s(&b, &m);
  if (method) *method = m;
  return r;
}

/* ------------------------------------------------------------------------ */
/* neumann.rs: bars, separation, point quadrature, sampled filaments         */
/* ------------------------------------------------------------------------ */

#define PARALLEL_SIN 1e-6
#define CLOUD_CAPACITY 256

typedef struct {
  v3 centre;
  v3 axes[3];
  double half[3];
} bar;

/* neumann.rs:67-73 */
static bar bar_new(const fho_filament *f) {
  bar b;
  b.centre = vmul(vadd(vload(f->start), vload(f->end)), 0.5);
  b.axes[0] = vload(f->dir);
  b.axes[1] = vload(f->wdir);
  b.axes[2] = vload(f->hdir);
  b.half[0] = 0.5 * f->length;
  b.half[1] = 0.5 * f->width;
  b.half[2] = 0.5 * f->height;
  return b;
}

static double circumradius(const bar *b) {
  return sqrt(b->half[0] * b->half[0] + b->half[1] * b->half[1] + b->half[2] * b->half[2]);
}

static double projected_half_extent(const bar *b, v3 n) {
  double s = 0.0;
  for (int k = 0; k < 3; ++k) s += b->half[k] * fabs(vdot(n, b->axes[k]));
  return s;
}

/* neumann.rs:97-122 */
static double separation(const bar *a, const bar *b) {
  const v3 delta = vsub(b->centre, a->centre);
  const double distance = vnorm(delta), reach = circumradius(a) + circumradius(b);
  if (distance > 8.0 * reach) return distance - reach;
  double best = -INFINITY;
#define CONSIDER(nv)                                                                     \
  do {                                                                                   \
    const v3 n_ = (nv);                                                                  \
    const double gap_ =                                                                  \
        fabs(vdot(n_, delta)) - projected_half_extent(a, n_) - projected_half_extent(b, n_); \
    best = fmax_r(best, gap_);                                                           \
  } while (0)
  for (int k = 0; k < 3; ++k) {
    CONSIDER(a->axes[k]);
    CONSIDER(b->axes[k]);
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const v3 n = vcross(a->axes[i], b->axes[j]);
      const double norm = vnorm(n);
      if (norm > 1e-9) CONSIDER(vdiv(n, norm));
    }
#undef CONSIDER
  return best;
}

/* neumann.rs:127-134; returns 0 for None. */
static int point_orders(const bar *b, double gap, int orders[3]) {
  for (int k = 0; k < 3; ++k) {
    orders[k] = order_for(g_point_thresholds, POINT_MAX_ORDER, b->half[k], gap);
    if (!orders[k]) return 0;
  }
  return orders[0] * orders[1] * orders[2] <= CLOUD_CAPACITY;
}

typedef struct {
  double x[CLOUD_CAPACITY], y[CLOUD_CAPACITY], z[CLOUD_CAPACITY], w[CLOUD_CAPACITY];
  int len;
} cloud;

/* neumann.rs:160-184 */
static void cloud_fill(cloud *c, const bar *b, const int orders[3]) {
  int i = 0;
  for (int il = 0; il < orders[0]; ++il) {
    double xl, wl;
    rule_on(orders[0], il, -b->half[0], b->half[0], &xl, &wl);
    for (int iw = 0; iw < orders[1]; ++iw) {
      double xw, ww;
      rule_on(orders[1], iw, -0.5, 0.5, &xw, &ww);
      for (int ih = 0; ih < orders[2]; ++ih) {
        double xh, wh;
        rule_on(orders[2], ih, -0.5, 0.5, &xh, &wh);
        const v3 p = vadd(vadd(vadd(b->centre, vmul(b->axes[0], xl)),
                               vmul(b->axes[1], 2.0 * b->half[1] * xw)),
                          vmul(b->axes[2], 2.0 * b->half[2] * xh));
        c->x[i] = p.x;
        c->y[i] = p.y;
        c->z[i] = p.z;
        c->w[i] = wl * ww * wh;
        ++i;
      }
    }
  }
  c->len = i;
  for (int pad = i; pad < ((i + 3) / 4) * 4; ++pad) {
    c->x[pad] = c->y[pad] = c->z[pad] = 1e150;
    c->w[pad] = 0.0;
  }
}

/* neumann.rs:237-266: four-lane inner sum, lanes reduced (l0 + l1) + (l2 + l3) */
static double point_sum_simd(const cloud *a, const cloud *b) {
  const int lanes = (b->len + 3) / 4;
  double total = 0.0;
  for (int i = 0; i < a->len; ++i) {
    const double px = a->x[i], py = a->y[i], pz = a->z[i];
    double inner[4] = {0.0, 0.0, 0.0, 0.0};
    for (int lane = 0; lane < lanes; ++lane)
      for (int k = 0; k < 4; ++k) {
        const int q = 4 * lane + k;
        const double dx = px - b->x[q], dy = py - b->y[q], dz = pz - b->z[q];
        inner[k] += b->w[q] / sqrt(dx * dx + dy * dy + dz * dz);
      }
    total += a->w[i] * ((inner[0] + inner[1]) + (inner[2] + inner[3]));
  }
  return total;
}

/* neumann.rs:280-325; returns 0 on a divergent coaxial sample. */
static int sampled_filaments(const bar *a, const bar *b, double gap, int parallel,
                             double *value_out, int *resolved_out) {
  int resolved = 1;
  int orders_a[2], orders_b[2];
  const double halves[4] = {a->half[1], a->half[2], b->half[1], b->half[2]};
  int *targets[4] = {&orders_a[0], &orders_a[1], &orders_b[0], &orders_b[1]};
  for (int q = 0; q < 4; ++q) {
    const int n = order_for(g_sample_thresholds, SAMPLE_MAX_ORDER, halves[q], gap);
    if (n) {
      *targets[q] = n;
    } else {
      resolved = 0;
      *targets[q] = SAMPLE_MAX_ORDER + (q >= 2 ? 1 : 0);
    }
  }
  v3 pa[81], pb[81];
  double wa[81], wb[81];
  int na = 0, nb = 0;
  for (int i = 0; i < orders_a[0]; ++i) {
    double xw, ww;
    rule_on(orders_a[0], i, -0.5, 0.5, &xw, &ww);
    for (int j = 0; j < orders_a[1]; ++j) {
      double xh, wh;
      rule_on(orders_a[1], j, -0.5, 0.5, &xh, &wh);
      const v3 off = vadd(vmul(a->axes[1], 2.0 * a->half[1] * xw),
                          vmul(a->axes[2], 2.0 * a->half[2] * xh));
      pa[na] = vadd(a->centre, off);
      wa[na++] = ww * wh;
    }
  }
  for (int i = 0; i < orders_b[0]; ++i) {
    double xw, ww;
    rule_on(orders_b[0], i, -0.5, 0.5, &xw, &ww);
    for (int j = 0; j < orders_b[1]; ++j) {
      double xh, wh;
      rule_on(orders_b[1], j, -0.5, 0.5, &xh, &wh);
      const v3 off = vadd(vmul(b->axes[1], 2.0 * b->half[1] * xw),
                          vmul(b->axes[2], 2.0 * b->half[2] * xh));
      pb[nb] = vadd(b->centre, off);
      wb[nb++] = ww * wh;
    }
  }
  const double la = 2.0 * a->half[0], lb = 2.0 * b->half[0];
  const v3 axis = a->axes[0];
  double value = 0.0;
  for (int i = 0; i < na; ++i)
    for (int j = 0; j < nb; ++j) {
      double kernel;
      if (parallel) {
        const v3 delta = vsub(pb[j], pa[i]);
        const double along = vdot(delta, axis);
        const double rho = vnorm(vsub(delta, vmul(axis, along)));
        if (!fho_lines_parallel(la, lb, along + 0.5 * (la - lb), rho, &kernel)) return 0;
      } else {
        const v3 ha = vmul(a->axes[0], a->half[0]), hb = vmul(b->axes[0], b->half[0]);
        if (!skew_v(vsub(pa[i], ha), vadd(pa[i], ha), vsub(pb[j], hb), vadd(pb[j], hb), 0.0,
                    &kernel))
          return 0;
      }
      value += wa[i] * wb[j] * kernel;
    }
  *value_out = value;
  *resolved_out = resolved;
  return 1;
}

/* ------------------------------------------------------------------------ */
/* inductance/mod.rs: per-pair dispatch                                      */
/* ------------------------------------------------------------------------ */

static int64_t total_key(double x) {
  int64_t a = (int64_t)bits_of(x);
  return a ^ (int64_t)(((uint64_t)(a >> 63)) >> 1);
}

/* mod.rs:231-254 */
static int canonical(const fho_filament *a, const fho_filament *b) {
  const double ka[11] = {a->start[0], a->start[1], a->start[2], a->end[0], a->end[1], a->end[2],
                         a->width,    a->height,   a->wdir[0],  a->wdir[1], a->wdir[2]};
  const double kb[11] = {b->start[0], b->start[1], b->start[2], b->end[0], b->end[1], b->end[2],
                         b->width,    b->height,   b->wdir[0],  b->wdir[1], b->wdir[2]};
  for (int i = 0; i < 11; ++i) {
    const int64_t x = total_key(ka[i]), y = total_key(kb[i]);
    if (x != y) return x < y ? -1 : 1;
  }
  return 0;
}

/* mod.rs:280-300 */
static int aligned_bars_of(const bar *a, const bar *b, aligned_bars *out) {
  const double ALIGNED = 1.0 - 1e-12;
  double width_b, height_b;
  if (fabs(vdot(a->axes[1], b->axes[1])) > ALIGNED) {
    width_b = b->half[1];
    height_b = b->half[2];
  } else if (fabs(vdot(a->axes[1], b->axes[2])) > ALIGNED) {
    width_b = b->half[2];
    height_b = b->half[1];
  } else {
    return 0;
  }
  const v3 delta = vsub(b->centre, a->centre);
  out->length[0] = 2.0 * a->half[0];
  out->length[1] = 2.0 * b->half[0];
  out->width[0] = 2.0 * a->half[1];
  out->width[1] = 2.0 * width_b;
  out->height[0] = 2.0 * a->half[2];
  out->height[1] = 2.0 * height_b;
  out->offset[0] = vdot(delta, a->axes[0]);
  out->offset[1] = vdot(delta, a->axes[1]);
  out->offset[2] = vdot(delta, a->axes[2]);
  return 1;
}

static double area_of(const fho_filament *f) { return f->width * f->height; }

/* mod.rs:303-369.  Returns FHO_OK, or FHO_E_KERNEL with code 1
 * (CoincidentSamples) or 2 (NotFinite). */
int fho_mutual(const fho_filament *a, const fho_filament *b, double *value_out, int *method_out,
               int *resolved_out) {
  ensure_tables();
  if (canonical(a, b) > 0) {
    const fho_filament *t = a;
    a = b;
    b = t;
  }
  const v3 da = vload(a->dir), db = vload(b->dir);
  const double cos_ = vdot(da, db);
  if (fabs(cos_) < 1e-14) {
    *value_out = 0.0;
    if (method_out) *method_out = FHO_METHOD_ORTHOGONAL;
    if (resolved_out) *resolved_out = 1;
    return FHO_OK;
  }
  const bar bar_a = bar_new(a), bar_b = bar_new(b);
  const int parallel = vnorm(vcross(da, db)) < PARALLEL_SIN;
  aligned_bars ab;
  const int aligned = parallel ? aligned_bars_of(&bar_a, &bar_b, &ab) : 0;
  const double gap = separation(&bar_a, &bar_b);
  int oa[3], ob[3];
  int have_orders = point_orders(&bar_a, gap, oa);
  if (have_orders) have_orders = point_orders(&bar_b, gap, ob);
  if (have_orders && aligned) {
    const int pairs = oa[0] * oa[1] * oa[2] * (ob[0] * ob[1] * ob[2]);
    if (pairs > 4096) have_orders = 0;
  }
  double integral;
  int method, resolved = 1;
  if (have_orders) {
    static __thread cloud ca, cb;
    cloud_fill(&ca, &bar_a, oa);
    cloud_fill(&cb, &bar_b, ob);
    integral = point_sum_simd(&ca, &cb) * cos_;
    method = FHO_METHOD_POINT_QUADRATURE;
  } else if (aligned) {
    const double in = aligned_dispatch_s(&ab, &method);
    const double scale = signum(cos_) / (area_of(a) * area_of(b));
    integral = in * scale;
  } else {
    double v;
    if (!sampled_filaments(&bar_a, &bar_b, gap, parallel, &v, &resolved))
      return fail(FHO_E_KERNEL, 1, 0, 0, 0.0);
    const double direction = parallel ? signum(cos_) : cos_;
    integral = v * direction;
    method = FHO_METHOD_SAMPLED_FILAMENTS;
  }
  const double value = 1.0e-7 * integral;
  if (!isfinite(value)) return fail(FHO_E_KERNEL, 2, 0, 0, value);
  *value_out = value;
  if (method_out) *method_out = method;
  if (resolved_out) *resolved_out = resolved;
  return FHO_OK;
}

/* mod.rs:196-202 */
double fho_self_inductance(const fho_filament *f) {
  ensure_tables();
  aligned_bars b = {{f->length, f->length}, {f->width, f->width}, {f->height, f->height},
                    {0.0, 0.0, 0.0}};
  int method;
  const double integral = aligned_dispatch_s(&b, &method);
  return 1.0e-7 * (integral * (1.0 / (area_of(f) * area_of(f))));
}

/* batch.rs:355-372: upper triangle evaluated, mirrored; rows in parallel. */
int fho_fill(const fho_filament *f, size_t n, double *L, int threads) {
  ensure_tables();
  int64_t err_i = -1, err_j = -1;
  int err_code = 0;
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    for (int64_t j = i; j < (int64_t)n; ++j) {
      double v;
      int m, r;
      const int st = fho_mutual(&f[i], &f[j], &v, &m, &r);
      if (st != FHO_OK) {
        const int code = g_error.code;
#pragma omp critical
        {
          if (err_i < 0 || i < err_i || (i == err_i && j < err_j)) {
            err_i = i;
            err_j = j;
            err_code = code;
          }
        }
        v = NAN;
      }
      L[i * n + j] = v;
      L[j * n + i] = v;
    }
  }
  if (err_i >= 0) return fail(FHO_E_KERNEL, err_code, err_i, err_j, 0.0);
  return FHO_OK;
}

int fho_fill_pairs(const fho_filament *f, size_t n, const int64_t *rows, const int64_t *cols,
                   size_t count, double *out, int *methods, int threads) {
  ensure_tables();
  int status = FHO_OK;
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
  for (int64_t k = 0; k < (int64_t)count; ++k) {
    if (rows[k] < 0 || cols[k] < 0 || (size_t)rows[k] >= n || (size_t)cols[k] >= n) {
      out[k] = NAN;
      continue;
    }
    double v;
    int m = -1, r;
    if (fho_mutual(&f[rows[k]], &f[cols[k]], &v, &m, &r) != FHO_OK) {
      v = NAN;
#pragma omp atomic write
      status = FHO_E_KERNEL;
    }
    out[k] = v;
    if (methods) methods[k] = m;
  }
  return status;
}

/* ------------------------------------------------------------------------ */
/* geometry.rs + filament.rs: segment frames and discretization              */
/* ------------------------------------------------------------------------ */

/* geometry.rs:45-54 */
static double scale_invariant_norm(v3 v) {
  if (!(isfinite(v.x) && isfinite(v.y) && isfinite(v.z))) return NAN;
  const double scale = fmax_r(fmax_r(fabs(v.x), fabs(v.y)), fabs(v.z));
  if (scale == 0.0) return 0.0;
  return vnorm(vdiv(v, scale)) * scale;
}

/* geometry.rs:223-276.  Sub-codes: 1 NonFiniteEndpoint, 2 NonPositiveDimension,
 * 3 NonPositiveConductivity, 4 ZeroLength, 5 InvalidWidthDirection. */
static int segment_basis(v3 a, v3 b, const fho_segment *s, v3 *l_out, v3 *w_out, v3 *h_out) {
  if (!(isfinite(a.x) && isfinite(a.y) && isfinite(a.z) && isfinite(b.x) && isfinite(b.y) &&
        isfinite(b.z)))
    return 1;
  if (!(isfinite(s->width) && s->width > 0.0)) return 2;
  if (!(isfinite(s->height) && s->height > 0.0)) return 2;
  if (!(isfinite(s->sigma) && s->sigma > 0.0)) return 3;
  const v3 span = vsub(b, a);
  const double norm = scale_invariant_norm(span);
  if (!(isfinite(norm) && norm > 0.0)) return 4;
  const v3 length = vdiv(span, norm);
  v3 hint;
  if (s->has_width_dir) {
    v3 dir = vload(s->width_dir);
    const double dn = scale_invariant_norm(dir);
    if (!(isfinite(dn) && dn > 0.0)) return 5;
    dir = vdiv(dir, dn);
    const v3 perpendicular = vsub(dir, vmul(length, vdot(dir, length)));
    if (vnorm(perpendicular) <= 1.0e-9) return 5;
    hint = perpendicular;
  } else {
    const v3 in_plane = vcross(V(0.0, 0.0, 1.0), length);
    hint = vnorm(in_plane) < 1.0e-6 ? V(1.0, 0.0, 0.0) : in_plane;
  }
  const v3 wv = vsub(hint, vmul(length, vdot(hint, length)));
  const v3 width = vdiv(wv, vnorm(wv));
  *l_out = length;
  *w_out = width;
  *h_out = vcross(length, width);
  return 0;
}

static void make_filament(fho_filament *f, v3 start, v3 end, double length, v3 l, v3 w, v3 h,
                          double width, double height, double sigma) {
  vstore(f->start, start);
  vstore(f->end, end);
  f->length = length;
  vstore(f->dir, l);
  vstore(f->wdir, w);
  vstore(f->hdir, h);
  f->width = width;
  f->height = height;
  f->sigma = sigma;
}

/* filament.rs:262-277 weights; 0 on invalid ratio / overflow */
static int graded_cells(double total, uint32_t count, double ratio, double *centre,
                        double *extent) {
  double sum = 0.0;
  for (uint32_t i = 0; i < count; ++i) {
    const uint32_t e = i < count - 1 - i ? i : count - 1 - i;
    extent[i] = powi_r(ratio, (int)e);
    if (!isfinite(extent[i])) return 0;
  }
  for (uint32_t i = 0; i < count; ++i) sum += extent[i];
  double lower = 0.0;
  for (uint32_t i = 0; i < count; ++i) {
    const double weight = extent[i];
    centre[i] = total * ((lower + 0.5 * weight) / sum - 0.5);
    lower += weight;
    extent[i] = total * (weight / sum);
  }
  return 1;
}

/* filament/grading.rs:11-55 */
static int grading_valid(v3 a, v3 b, double seg_len, const fho_segment *s, v3 wdir, v3 hdir,
                         const double *wc, const double *we, uint32_t nw, const double *hc,
                         const double *he, uint32_t nh) {
  const double shape_scale = fmax_r(fmax_r(seg_len, s->width), s->height);
  const v3 aa = vabs(a), ab = vabs(b);
  const v3 sup = V(fmax_r(aa.x, ab.x), fmax_r(aa.y, ab.y), fmax_r(aa.z, ab.z));
  const v3 envelope =
      vadd(vadd(sup, vmul(vabs(wdir), 0.5 * s->width)), vmul(vabs(hdir), 0.5 * s->height));
  for (int axis = 0; axis < 2; ++axis) {
    const v3 direction = axis == 0 ? wdir : hdir;
    const double *c = axis == 0 ? wc : hc;
    const double *e = axis == 0 ? we : he;
    const uint32_t n = axis == 0 ? nw : nh;
    const double coordinate_scale = vdot(vabs(direction), envelope);
    const double minimum =
        fmax_r(1e-7 * shape_scale, 32.0 * 2.220446049250313e-16 * coordinate_scale);
    for (uint32_t i = 0; i < n; ++i) {
      const double size = e[i], centre = c[i];
      if (!(isfinite(size) && size > 0.0 && size >= minimum && isfinite(centre) &&
            centre - 0.5 * size < centre && centre + 0.5 * size > centre))
        return 0;
      if (i + 1 < n && !(c[i] < c[i + 1])) return 0;
    }
  }
  return 1;
}

/* Discretize error sub-codes: 10 ZeroSubdivision, 11 Overflow, 12 InvalidRatio,
 * 13 GradingOverflow, 14 UnusableGrading, 1..5 segment errors. */
static int discretize_segment(v3 a, v3 b, const fho_segment *s, const fho_grid *g,
                              fho_filament *out) {
  if (g->graded) {
    if (!(isfinite(g->width_ratio) && g->width_ratio >= 1.0)) return 12;
    if (!(isfinite(g->height_ratio) && g->height_ratio >= 1.0)) return 12;
  }
  if (g->nw == 0 || g->nh == 0) return 10;
  if ((uint64_t)g->nw * (uint64_t)g->nh > (1ULL << 31)) return 11;
  v3 l, w, h;
  const int be = segment_basis(a, b, s, &l, &w, &h);
  if (be) return be;
  const double length = vnorm(vsub(b, a));
  if (g->graded) {
    double *wc = malloc(sizeof(double) * 2 * (g->nw + g->nh));
    if (!wc) return 11;
    double *we = wc + g->nw, *hc = we + g->nw, *he = hc + g->nh;
    if (!graded_cells(s->width, g->nw, g->width_ratio, wc, we) ||
        !graded_cells(s->height, g->nh, g->height_ratio, hc, he)) {
      free(wc);
      return 13;
    }
    if (!grading_valid(a, b, length, s, w, h, wc, we, g->nw, hc, he, g->nh)) {
      free(wc);
      return 14;
    }
    if (!(g->width_ratio == 1.0 && g->height_ratio == 1.0)) {
      size_t k = 0;
      for (uint32_t j = 0; j < g->nh; ++j) {
        const v3 along_height = vmul(h, hc[j]);
        for (uint32_t i = 0; i < g->nw; ++i) {
          const v3 offset = vadd(vmul(w, wc[i]), along_height);
          make_filament(&out[k++], vadd(a, offset), vadd(b, offset), length, l, w, h, we[i],
                        he[j], s->sigma);
        }
      }
      free(wc);
      return 0;
    }
    free(wc);
  }
  /* filament.rs:320-359 */
  const double fw = s->width / (double)g->nw, fh = s->height / (double)g->nh;
  size_t k = 0;
  for (uint32_t j = 0; j < g->nh; ++j) {
    const double fj = ((double)j + 0.5) / (double)g->nh - 0.5;
    const v3 along_height = vmul(h, fj * s->height);
    for (uint32_t i = 0; i < g->nw; ++i) {
      const double fi = ((double)i + 0.5) / (double)g->nw - 0.5;
      const v3 offset = vadd(vmul(w, fi * s->width), along_height);
      make_filament(&out[k++], vadd(a, offset), vadd(b, offset), length, l, w, h, fw, fh,
                    s->sigma);
    }
  }
  return 0;
}

static int validate_geometry(const double *nodes, size_t n_nodes, const fho_segment *segments,
                             size_t n_segments) {
  for (size_t i = 0; i < n_nodes; ++i)
    if (!(isfinite(nodes[3 * i]) && isfinite(nodes[3 * i + 1]) && isfinite(nodes[3 * i + 2])))
      return fail(FHO_E_GEOMETRY, 1, (int64_t)i, 0, 0.0);
  for (size_t s = 0; s < n_segments; ++s) {
    if (segments[s].a >= n_nodes || segments[s].b >= n_nodes)
      return fail(FHO_E_GEOMETRY, 2, (int64_t)s, 0, 0.0);
    v3 l, w, h;
    const int be = segment_basis(vload(nodes + 3 * segments[s].a),
                                 vload(nodes + 3 * segments[s].b), &segments[s], &l, &w, &h);
    if (be) return fail(FHO_E_GEOMETRY, 10 + be, (int64_t)s, 0, 0.0);
  }
  return FHO_OK;
}

int fho_discretize_geometry(const double *nodes, size_t n_nodes, const fho_segment *segments,
                            size_t n_segments, const fho_grid *grids, fho_filament *out,
                            size_t capacity, size_t *n_out) {
  const int vg = validate_geometry(nodes, n_nodes, segments, n_segments);
  if (vg) return vg;
  size_t total = 0;
  for (size_t s = 0; s < n_segments; ++s) total += (size_t)grids[s].nw * grids[s].nh;
  *n_out = total;
  if (!out) return FHO_OK;
  if (capacity < total) return fail(FHO_E_INVALID_ARGUMENT, 0, 0, 0, 0.0);
  size_t k = 0;
  for (size_t s = 0; s < n_segments; ++s) {
    const int e = discretize_segment(vload(nodes + 3 * segments[s].a),
                                     vload(nodes + 3 * segments[s].b), &segments[s], &grids[s],
                                     out + k);
    if (e) return fail(e >= 12 && e <= 12 ? FHO_E_INVALID_GRADING : FHO_E_DISCRETIZE, e,
                       (int64_t)s, 0, 0.0);
    k += (size_t)grids[s].nw * grids[s].nh;
  }
  return FHO_OK;
}

/* ------------------------------------------------------------------------ */
/* mesh.rs: loop basis                                                       */
/* ------------------------------------------------------------------------ */

typedef struct {
  int64_t parent, segment;
  int forwards;
} parent_link;

/* Mesh error sub-codes: 1 NoPorts, 2 UnknownPortNode, 3 DegeneratePort,
 * 4 PortNotConnected, 5 SegmentCountMismatch, 6 EmptySegment. */
int fho_mesh_build(size_t n_nodes, const fho_segment *segments, size_t n_segments,
                   const size_t *per_segment, const fho_port *ports, size_t n_ports,
                   int64_t *row_ptr, int64_t *branch, double *sign, size_t *loops_out,
                   size_t *internal_out, size_t *nnz_out) {
  for (size_t s = 0; s < n_segments; ++s)
    if (per_segment[s] == 0) return fail(FHO_E_MESH, 6, (int64_t)s, 0, 0.0);
  if (n_ports == 0) return fail(FHO_E_MESH, 1, 0, 0, 0.0);
  int64_t *first = malloc(sizeof(int64_t) * (n_segments + 1));
  /* adjacency in segment order (mesh.rs:290-296) */
  int64_t *adj_ptr = calloc(n_nodes + 1, sizeof(int64_t));
  int64_t *adj = malloc(sizeof(int64_t) * 3 * (2 * n_segments + 1));
  parent_link *parent = malloc(sizeof(parent_link) * (n_nodes + 1));
  int64_t *depth = calloc(n_nodes + 1, sizeof(int64_t));
  int64_t *component = malloc(sizeof(int64_t) * (n_nodes + 1));
  char *in_tree = calloc(n_segments + 1, 1);
  int64_t *queue = malloc(sizeof(int64_t) * (n_nodes + 1));
  int64_t *fill = calloc(n_nodes + 1, sizeof(int64_t));
  /* a path has at most n_nodes entries */
  int64_t *path_seg = malloc(sizeof(int64_t) * (n_nodes + 1));
  double *path_sign = malloc(sizeof(double) * (n_nodes + 1));
  int64_t *down_seg = malloc(sizeof(int64_t) * (n_nodes + 1));
  double *down_sign = malloc(sizeof(double) * (n_nodes + 1));
  int status = FHO_OK;
  if (!first || !adj_ptr || !adj || !parent || !depth || !component || !in_tree || !queue ||
      !fill || !path_seg || !path_sign || !down_seg || !down_sign) {
    status = fail(FHO_E_NO_MEMORY, 0, 0, 0, 0.0);
    goto done;
  }
  first[0] = 0;
  for (size_t s = 0; s < n_segments; ++s) first[s + 1] = first[s] + (int64_t)per_segment[s];
  for (size_t s = 0; s < n_segments; ++s) {
    adj_ptr[segments[s].a + 1]++;
    adj_ptr[segments[s].b + 1]++;
  }
  for (size_t v = 0; v < n_nodes; ++v) adj_ptr[v + 1] += adj_ptr[v];
  for (size_t s = 0; s < n_segments; ++s) {
    const uint32_t a = segments[s].a, b = segments[s].b;
    int64_t *e = adj + 3 * (adj_ptr[a] + fill[a]++);
    e[0] = b;
    e[1] = (int64_t)s;
    e[2] = 1;
    e = adj + 3 * (adj_ptr[b] + fill[b]++);
    e[0] = a;
    e[1] = (int64_t)s;
    e[2] = 0;
  }
  for (size_t v = 0; v < n_nodes; ++v) {
    parent[v].parent = -1;
    component[v] = -1;
  }
  int64_t components = 0;
  for (size_t root = 0; root < n_nodes; ++root) {
    if (component[root] >= 0 || adj_ptr[root + 1] == adj_ptr[root]) continue;
    component[root] = components;
    size_t qh = 0, qt = 0;
    queue[qt++] = (int64_t)root;
    while (qh < qt) {
      const int64_t node = queue[qh++];
      for (int64_t k = adj_ptr[node]; k < adj_ptr[node + 1]; ++k) {
        const int64_t next = adj[3 * k], seg = adj[3 * k + 1];
        if (component[next] < 0) {
          component[next] = components;
          parent[next].parent = node;
          parent[next].segment = seg;
          parent[next].forwards = (int)adj[3 * k + 2];
          depth[next] = depth[node] + 1;
          in_tree[seg] = 1;
          queue[qt++] = next;
        }
      }
    }
    components++;
  }
  /* mesh.rs:344-360 path(from, to) -> (path_seg, path_sign, count) */
#define PATH(from, to, count_out)                                            \
  do {                                                                       \
    int64_t x_ = (from), y_ = (to), nu_ = 0, nd_ = 0;                         \
    while (x_ != y_) {                                                       \
      if (depth[x_] >= depth[y_]) {                                          \
        path_seg[nu_] = parent[x_].segment;                                  \
        path_sign[nu_++] = parent[x_].forwards ? -1.0 : 1.0;                 \
        x_ = parent[x_].parent;                                              \
      } else {                                                               \
        down_seg[nd_] = parent[y_].segment;                                  \
        down_sign[nd_++] = parent[y_].forwards ? 1.0 : -1.0;                 \
        y_ = parent[y_].parent;                                              \
      }                                                                      \
    }                                                                        \
    for (int64_t q_ = nd_ - 1; q_ >= 0; --q_) {                               \
      path_seg[nu_] = down_seg[q_];                                          \
      path_sign[nu_++] = down_sign[q_];                                      \
    }                                                                        \
    (count_out) = nu_;                                                       \
  } while (0)
  /* Validate ports first (mesh.rs:215-240) so that errors precede output. */
  for (size_t p = 0; p < n_ports; ++p) {
    const uint32_t pos = ports[p].positive, neg = ports[p].negative;
    if (pos >= n_nodes) {
      status = fail(FHO_E_MESH, 2, (int64_t)p, pos, 0.0);
      goto done;
    }
    if (neg >= n_nodes) {
      status = fail(FHO_E_MESH, 2, (int64_t)p, neg, 0.0);
      goto done;
    }
    if (pos == neg) {
      status = fail(FHO_E_MESH, 3, (int64_t)p, pos, 0.0);
      goto done;
    }
    if (!(component[pos] >= 0 && component[neg] >= 0 && component[pos] == component[neg])) {
      status = fail(FHO_E_MESH, 4, (int64_t)p, pos, (double)neg);
      goto done;
    }
  }
  {
    int64_t loops = 0, nnz = 0, internal;
    /* bundle loops */
    for (size_t s = 0; s < n_segments; ++s)
      for (size_t k = 1; k < per_segment[s]; ++k) {
        if (row_ptr) {
          row_ptr[loops] = nnz;
          branch[nnz] = first[s] + (int64_t)k;
          sign[nnz] = 1.0;
          branch[nnz + 1] = first[s];
          sign[nnz + 1] = -1.0;
        }
        nnz += 2;
        loops++;
      }
    /* conductor loops */
    for (size_t s = 0; s < n_segments; ++s) {
      if (in_tree[s]) continue;
      int64_t count;
      PATH(segments[s].b, segments[s].a, count);
      if (row_ptr) {
        row_ptr[loops] = nnz;
        branch[nnz] = first[s];
        sign[nnz] = 1.0;
        for (int64_t q = 0; q < count; ++q) {
          branch[nnz + 1 + q] = first[path_seg[q]];
          sign[nnz + 1 + q] = path_sign[q];
        }
      }
      nnz += 1 + count;
      loops++;
    }
    internal = loops;
    for (size_t p = 0; p < n_ports; ++p) {
      int64_t count;
      PATH(ports[p].positive, ports[p].negative, count);
      if (row_ptr) {
        row_ptr[loops] = nnz;
        for (int64_t q = 0; q < count; ++q) {
          branch[nnz + q] = first[path_seg[q]];
          sign[nnz + q] = path_sign[q];
        }
      }
      nnz += count;
      loops++;
    }
    if (row_ptr) row_ptr[loops] = nnz;
    *loops_out = (size_t)loops;
    *internal_out = (size_t)internal;
    *nnz_out = (size_t)nnz;
  }
#undef PATH
done:
  free(first);
  free(adj_ptr);
  free(adj);
  free(parent);
  free(depth);
  free(component);
  free(in_tree);
  free(queue);
  free(fill);
  free(path_seg);
  free(path_sign);
  free(down_seg);
  free(down_sign);
  return status;
}

/* ------------------------------------------------------------------------ */
/* solve.rs:1172-1200 congruence                                             */
/* ------------------------------------------------------------------------ */

int fho_congruence(size_t loops, const int64_t *row_ptr, const int64_t *branch,
                   const double *sign, size_t n, const double *L, const double *r, double *out,
                   int threads) {
  if (loops == 0 || n == 0) {
    memset(out, 0, sizeof(double) * loops * loops);
    return FHO_OK;
  }
  if (threads <= 0) threads = omp_get_max_threads();
  int status = FHO_OK;
#pragma omp parallel num_threads(threads)
  {
    double *product = malloc(sizeof(double) * n);
    if (!product) {
#pragma omp atomic write
      status = FHO_E_NO_MEMORY;
    }
#pragma omp for schedule(dynamic, 4)
    for (int64_t i = 0; i < (int64_t)loops; ++i) {
      if (!product) continue;
      for (size_t k = 0; k < n; ++k) product[k] = 0.0;
      for (int64_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
        const int64_t b = branch[e];
        const double s = sign[e];
        if (L) {
          const double *col = L + (size_t)b * n; /* symmetric: column b = row b */
          for (size_t k = 0; k < n; ++k) product[k] += s * col[k];
        } else {
          product[b] += s * r[b];
        }
      }
      for (int64_t j = i; j < (int64_t)loops; ++j) {
        double acc = 0.0;
        for (int64_t e = row_ptr[j]; e < row_ptr[j + 1]; ++e) acc += sign[e] * product[branch[e]];
        out[i * loops + j] = acc;
        out[j * loops + i] = acc;
      }
    }
    free(product);
  }
  return status;
}

/* ------------------------------------------------------------------------ */
/* dense.rs: blocked right-looking LU with partial pivoting                  */
/* ------------------------------------------------------------------------ */

#define PANEL 32
#define PARALLEL_THRESHOLD (1 << 14)

typedef struct {
  double re, im;
} cplx;

static inline cplx cmul(cplx a, cplx b) {
  cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return r;
}
/* num-complex 0.4 Div: (ac + bd)/n + i (bc - ad)/n, n = c^2 + d^2 */
static inline cplx cdiv(cplx a, cplx b) {
  const double n = b.re * b.re + b.im * b.im;
  cplx r = {(a.re * b.re + a.im * b.im) / n, (a.im * b.re - a.re * b.im) / n};
  return r;
}
static inline cplx csub(cplx a, cplx b) {
  cplx r = {a.re - b.re, a.im - b.im};
  return r;
}
static inline double cnorm1(cplx a) { return fabs(a.re) + fabs(a.im); }

#define DEFINE_LU(NAME, T, MUL, DIV, SUB, NORM1, ONE, ZERO_TEST)                             \
  static void NAME##_subtract_scaled(T *y, T factor, const T *x, size_t len) {              \
    for (size_t i = 0; i < len; ++i) y[i] = SUB(y[i], MUL(factor, x[i]));                  \
  }                                                                                         \
  static int NAME##_factorize_panel(T *panel, size_t order, size_t start, size_t *pivots,   \
                                    size_t width) {                                         \
    for (size_t c = 0; c < width; ++c) {                                                    \
      const size_t diagonal = start + c;                                                    \
      T *column = panel + c * order;                                                        \
      size_t pivot = diagonal;                                                              \
      double largest = NORM1(column[diagonal]);                                             \
      for (size_t row = diagonal + 1; row < order; ++row) {                                 \
        const double magnitude = NORM1(column[row]);                                        \
        if (magnitude > largest) {                                                          \
          pivot = row;                                                                      \
          largest = magnitude;                                                              \
        }                                                                                   \
      }                                                                                     \
      if (largest == 0.0) return 0;                                                         \
      pivots[c] = pivot;                                                                    \
      if (pivot != diagonal)                                                                \
        for (size_t k = 0; k < width; ++k) {                                                \
          T t = panel[k * order + diagonal];                                                \
          panel[k * order + diagonal] = panel[k * order + pivot];                           \
          panel[k * order + pivot] = t;                                                     \
        }                                                                                   \
      const T inverse = DIV(ONE, column[diagonal]);                                         \
      for (size_t r = diagonal + 1; r < order; ++r) column[r] = MUL(column[r], inverse);    \
      for (size_t k = c + 1; k < width; ++k) {                                              \
        T *other = panel + k * order;                                                       \
        NAME##_subtract_scaled(other + diagonal + 1, other[diagonal], column + diagonal + 1, \
                               order - diagonal - 1);                                       \
      }                                                                                     \
    }                                                                                       \
    return 1;                                                                               \
  }                                                                                         \
  static int NAME##_factorize(T *a, size_t order, size_t *pivots, int threads) {            \
    for (size_t start = 0; start < order; start += PANEL) {                                 \
      const size_t width = PANEL < order - start ? PANEL : order - start;                   \
      T *panel = a + start * order;                                                         \
      if (!NAME##_factorize_panel(panel, order, start, pivots + start, width)) return 0;   \
      const size_t *pv = pivots + start;                                                    \
      for (size_t col = 0; col < start; ++col) {                                            \
        T *column = a + col * order;                                                        \
        for (size_t o = 0; o < width; ++o) {                                                \
          T t = column[start + o];                                                          \
          column[start + o] = column[pv[o]];                                                \
          column[pv[o]] = t;                                                                \
        }                                                                                   \
      }                                                                                     \
      const size_t first_right = start + width;                                             \
      const int64_t nright = (int64_t)(order - first_right);                                \
      const int par = (order - start) * (size_t)nright >= PARALLEL_THRESHOLD && threads > 1; \
      _Pragma("omp parallel for schedule(static) num_threads(threads) if (par)")            \
      for (int64_t q = 0; q < nright; ++q) {                                                \
        T *column = a + (first_right + (size_t)q) * order;                                  \
        for (size_t o = 0; o < width; ++o) {                                                \
          T t = column[start + o];                                                          \
          column[start + o] = column[pv[o]];                                                \
          column[pv[o]] = t;                                                                \
        }                                                                                   \
        for (size_t c = 0; c < width; ++c) {                                                \
          const T *l = panel + c * order + start + c + 1;                                   \
          NAME##_subtract_scaled(column + start + c + 1, column[start + c], l,              \
                                 order - start - c - 1);                                    \
        }                                                                                   \
      }                                                                                     \
    }                                                                                       \
    return 1;                                                                               \
  }                                                                                         \
  static int NAME##_solve(T *a, size_t order, T *b, size_t nrhs, int threads) {             \
    if (order == 0) return FHO_OK;                                                          \
    size_t *pivots = malloc(sizeof(size_t) * order);                                        \
    if (!pivots) return FHO_E_NO_MEMORY;                                                    \
    if (!NAME##_factorize(a, order, pivots, threads)) {                                     \
      free(pivots);                                                                         \
      return FHO_E_SINGULAR;                                                                \
    }                                                                                       \
    const int par = order * order >= PARALLEL_THRESHOLD && threads > 1 && nrhs > 1;         \
    _Pragma("omp parallel for schedule(static) num_threads(threads) if (par)")              \
    for (int64_t q = 0; q < (int64_t)nrhs; ++q) {                                           \
      T *rhs = b + (size_t)q * order;                                                       \
      for (size_t row = 0; row < order; ++row) {                                            \
        T t = rhs[row];                                                                     \
        rhs[row] = rhs[pivots[row]];                                                        \
        rhs[pivots[row]] = t;                                                               \
      }                                                                                     \
      for (size_t c = 0; c < order; ++c)                                                    \
        NAME##_subtract_scaled(rhs + c + 1, rhs[c], a + c * order + c + 1, order - c - 1); \
      for (size_t c = order; c-- > 0;) {                                                    \
        rhs[c] = DIV(rhs[c], a[c * order + c]);                                             \
        NAME##_subtract_scaled(rhs, rhs[c], a + c * order, c);                              \
      }                                                                                     \
    }                                                                                       \
    free(pivots);                                                                           \
    return FHO_OK;                                                                          \
  }

static inline double rmul(double a, double b) { return a * b; }
static inline double rdiv(double a, double b) { return a / b; }
static inline double rsub(double a, double b) { return a - b; }
static const cplx CONE = {1.0, 0.0};

DEFINE_LU(clu, cplx, cmul, cdiv, csub, cnorm1, CONE, 0)
DEFINE_LU(rlu, double, rmul, rdiv, rsub, fabs, 1.0, 0)

int fho_lu_solve_complex(double *a, size_t order, double *b, size_t nrhs, int threads) {
  if (threads <= 0) threads = omp_get_max_threads();
  return clu_solve((cplx *)a, order, (cplx *)b, nrhs, threads);
}

int fho_lu_solve_real(double *a, size_t order, double *b, size_t nrhs, int threads) {
  if (threads <= 0) threads = omp_get_max_threads();
  return rlu_solve(a, order, b, nrhs, threads);
}

/* ------------------------------------------------------------------------ */
/* solve.rs:680-894: assembled system, impedance, sweep                      */
/* ------------------------------------------------------------------------ */

struct fho_system {
  size_t counts[6];
  size_t n, m, p;
  fho_filament *filaments;
  double *resistances;
  double *L;              /* n x n */
  double *r_mesh, *l_mesh; /* loops x loops row-major */
};

void fho_system_destroy(fho_system *s) {
  if (!s) return;
  free(s->filaments);
  free(s->resistances);
  free(s->L);
  free(s->r_mesh);
  free(s->l_mesh);
  free(s);
}

int fho_system_create(const double *nodes, size_t n_nodes, const fho_segment *segments,
                      size_t n_segments, const fho_grid *grids, const fho_port *ports,
                      size_t n_ports, int threads, fho_system **out) {
  *out = NULL;
  size_t n;
  int st = fho_discretize_geometry(nodes, n_nodes, segments, n_segments, grids, NULL, 0, &n);
  if (st) return st;
  fho_system *s = calloc(1, sizeof(fho_system));
  if (!s) return fail(FHO_E_NO_MEMORY, 0, 0, 0, 0.0);
  s->filaments = malloc(sizeof(fho_filament) * (n ? n : 1));
  st = fho_discretize_geometry(nodes, n_nodes, segments, n_segments, grids, s->filaments, n, &n);
  if (st) goto fail_;
  size_t *per = malloc(sizeof(size_t) * (n_segments ? n_segments : 1));
  for (size_t q = 0; q < n_segments; ++q) per[q] = (size_t)grids[q].nw * grids[q].nh;
  size_t loops, internal, nnz;
  st = fho_mesh_build(n_nodes, segments, n_segments, per, ports, n_ports, NULL, NULL, NULL,
                      &loops, &internal, &nnz);
  if (st) {
    free(per);
    goto fail_;
  }
  int64_t *row_ptr = malloc(sizeof(int64_t) * (loops + 1));
  int64_t *branch = malloc(sizeof(int64_t) * (nnz ? nnz : 1));
  double *sign = malloc(sizeof(double) * (nnz ? nnz : 1));
  fho_mesh_build(n_nodes, segments, n_segments, per, ports, n_ports, row_ptr, branch, sign,
                 &loops, &internal, &nnz);
  free(per);
  s->n = n;
  s->m = internal;
  s->p = loops - internal;
  s->resistances = malloc(sizeof(double) * (n ? n : 1));
  for (size_t i = 0; i < n; ++i) {
    const fho_filament *f = &s->filaments[i];
    s->resistances[i] = f->length / (f->sigma * area_of(f)); /* solve.rs:941-943 */
  }
  s->L = malloc(sizeof(double) * (n ? n * n : 1));
  s->r_mesh = malloc(sizeof(double) * (loops ? loops * loops : 1));
  s->l_mesh = malloc(sizeof(double) * (loops ? loops * loops : 1));
  if (!s->L || !s->r_mesh || !s->l_mesh) {
    st = fail(FHO_E_NO_MEMORY, 0, 0, 0, 0.0);
    free(row_ptr);
    free(branch);
    free(sign);
    goto fail_;
  }
  st = fho_fill(s->filaments, n, s->L, threads);
  if (!st) st = fho_congruence(loops, row_ptr, branch, sign, n, NULL, s->resistances, s->r_mesh,
                               threads);
  if (!st) st = fho_congruence(loops, row_ptr, branch, sign, n, s->L, NULL, s->l_mesh, threads);
  free(row_ptr);
  free(branch);
  free(sign);
  if (st) goto fail_;
  s->counts[0] = n_nodes;
  s->counts[1] = n_segments;
  s->counts[2] = n;
  s->counts[3] = loops;
  s->counts[4] = internal;
  s->counts[5] = loops - internal;
  *out = s;
  return FHO_OK;
fail_:
  fho_system_destroy(s);
  return st;
}

void fho_system_counts(const fho_system *s, size_t counts[6]) {
  memcpy(counts, s->counts, sizeof(s->counts));
}
const double *fho_system_inductance(const fho_system *s) { return s->L; }
const fho_filament *fho_system_filaments(const fho_system *s) { return s->filaments; }

int fho_system_blocks(const fho_system *s, double *r_ee, double *l_ee, double *r_ep,
                      double *l_ep, double *r_pp, double *l_pp) {
  const size_t m = s->m, p = s->p, loops = m + p;
  for (size_t i = 0; i < loops; ++i)
    for (size_t j = 0; j < loops; ++j) {
      const double r = s->r_mesh[i * loops + j], l = s->l_mesh[i * loops + j];
      if (i < m && j < m) {
        if (r_ee) r_ee[i * m + j] = r;
        if (l_ee) l_ee[i * m + j] = l;
      } else if (i < m && j >= m) {
        if (r_ep) r_ep[i * p + (j - m)] = r;
        if (l_ep) l_ep[i * p + (j - m)] = l;
      } else if (i >= m && j >= m) {
        if (r_pp) r_pp[(i - m) * p + (j - m)] = r;
        if (l_pp) l_pp[(i - m) * p + (j - m)] = l;
      }
    }
  return FHO_OK;
}

/* solve.rs:806-830 with schur_complement (solve.rs:1154-1164).  z: p x p
 * row-major interleaved complex. */
static int impedance_at(const fho_system *s, double frequency, double *z, int threads) {
  const size_t m = s->m, p = s->p, loops = m + p;
  if (!(isfinite(frequency) && frequency >= 0.0))
    return fail(FHO_E_INVALID_FREQUENCY, 0, 0, 0, frequency);
  if (frequency == 0.0) {
    /* real arithmetic branch */
    if (m == 0) {
      for (size_t i = 0; i < p; ++i)
        for (size_t j = 0; j < p; ++j) {
          z[2 * (i * p + j)] = s->r_mesh[(m + i) * loops + m + j];
          z[2 * (i * p + j) + 1] = 0.0;
        }
      return FHO_OK;
    }
    double *ee = malloc(sizeof(double) * m * m), *ep = malloc(sizeof(double) * m * p);
    if (!ee || !ep) {
      free(ee);
      free(ep);
      return fail(FHO_E_NO_MEMORY, 0, 0, 0, 0.0);
    }
    for (size_t c = 0; c < m; ++c)
      for (size_t r = 0; r < m; ++r) ee[c * m + r] = s->r_mesh[r * loops + c];
    for (size_t c = 0; c < p; ++c)
      for (size_t r = 0; r < m; ++r) ep[c * m + r] = s->r_mesh[r * loops + m + c];
    double *solved = malloc(sizeof(double) * m * p);
    memcpy(solved, ep, sizeof(double) * m * p);
    const int st = fho_lu_solve_real(ee, m, solved, p, threads);
    if (st) {
      free(ee);
      free(ep);
      free(solved);
      return fail(FHO_E_SINGULAR, 0, 0, 0, frequency);
    }
    for (size_t i = 0; i < p; ++i)
      for (size_t j = 0; j < p; ++j) {
        double acc = ep[i * m + 0] * solved[j * m + 0];
        for (size_t k = 1; k < m; ++k) acc = acc + ep[i * m + k] * solved[j * m + k];
        z[2 * (i * p + j)] = s->r_mesh[(m + i) * loops + m + j] - acc;
        z[2 * (i * p + j) + 1] = 0.0;
      }
    free(ee);
    free(ep);
    free(solved);
  } else {
    const double omega = 6.283185307179586 * frequency;
    if (m == 0) {
      for (size_t i = 0; i < p; ++i)
        for (size_t j = 0; j < p; ++j) {
          z[2 * (i * p + j)] = s->r_mesh[(m + i) * loops + m + j];
          z[2 * (i * p + j) + 1] = omega * s->l_mesh[(m + i) * loops + m + j];
        }
    } else {
      cplx *ee = malloc(sizeof(cplx) * m * m), *ep = malloc(sizeof(cplx) * m * p);
      cplx *solved = malloc(sizeof(cplx) * m * p);
      if (!ee || !ep || !solved) {
        free(ee);
        free(ep);
        free(solved);
        return fail(FHO_E_NO_MEMORY, 0, 0, 0, 0.0);
      }
      for (size_t c = 0; c < m; ++c)
        for (size_t r = 0; r < m; ++r) {
          cplx v = {s->r_mesh[r * loops + c], omega * s->l_mesh[r * loops + c]};
          ee[c * m + r] = v;
        }
      for (size_t c = 0; c < p; ++c)
        for (size_t r = 0; r < m; ++r) {
          cplx v = {s->r_mesh[r * loops + m + c], omega * s->l_mesh[r * loops + m + c]};
          ep[c * m + r] = v;
        }
      memcpy(solved, ep, sizeof(cplx) * m * p);
      const int st = clu_solve(ee, m, solved, p, threads);
      if (st) {
        free(ee);
        free(ep);
        free(solved);
        return fail(FHO_E_SINGULAR, 0, 0, 0, frequency);
      }
      for (size_t i = 0; i < p; ++i)
        for (size_t j = 0; j < p; ++j) {
          cplx acc = cmul(ep[i * m + 0], solved[j * m + 0]);
          for (size_t k = 1; k < m; ++k) {
            const cplx t = cmul(ep[i * m + k], solved[j * m + k]);
            acc.re = acc.re + t.re;
            acc.im = acc.im + t.im;
          }
          z[2 * (i * p + j)] = s->r_mesh[(m + i) * loops + m + j] - acc.re;
          z[2 * (i * p + j) + 1] = omega * s->l_mesh[(m + i) * loops + m + j] - acc.im;
        }
      free(ee);
      free(ep);
      free(solved);
    }
  }
  for (size_t q = 0; q < 2 * p * p; ++q)
    if (!isfinite(z[q])) return fail(FHO_E_SINGULAR, 0, 0, 0, frequency);
  return FHO_OK;
}

/* solve.rs:1039-1079: validate all, then chunks of concurrent frequencies
 * within the 2 GiB budget; the first failure in list order is reported. */
int fho_system_sweep(const fho_system *s, const double *freqs, size_t K, double *z_out,
                     int threads) {
  for (size_t k = 0; k < K; ++k)
    if (!(isfinite(freqs[k]) && freqs[k] >= 0.0))
      return fail(FHO_E_INVALID_FREQUENCY, 0, (int64_t)k, 0, freqs[k]);
  if (threads <= 0) threads = omp_get_max_threads();
  const size_t m = s->m, p = s->p;
  size_t bytes = 2 * m * (m + p) * 16;
  size_t concurrent = ((size_t)2 << 30) / (bytes ? bytes : 1);
  if (concurrent < 1) concurrent = 1;
  int *status = calloc(K ? K : 1, sizeof(int));
  fho_error *errs = calloc(K ? K : 1, sizeof(fho_error));
  for (size_t c0 = 0; c0 < K; c0 += concurrent) {
    const size_t c1 = c0 + concurrent < K ? c0 + concurrent : K;
    const int64_t count = (int64_t)(c1 - c0);
    /* Frequencies concurrently, each LU then single-threaded, when there
     * are at least as many frequencies as threads; otherwise one at a time
     * with the LU's trailing update in parallel. */
    if (count >= threads || threads == 1) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
      for (int64_t q = 0; q < count; ++q) {
        const size_t k = c0 + (size_t)q;
        status[k] = impedance_at(s, freqs[k], z_out + 2 * p * p * k, 1);
        if (status[k]) errs[k] = g_error;
      }
    } else {
      for (int64_t q = 0; q < count; ++q) {
        const size_t k = c0 + (size_t)q;
        status[k] = impedance_at(s, freqs[k], z_out + 2 * p * p * k, threads);
        if (status[k]) errs[k] = g_error;
      }
    }
    for (size_t k = c0; k < c1; ++k)
      if (status[k]) {
        g_error = errs[k];
        g_error.index0 = (int64_t)k;
        const int st = status[k];
        free(status);
        free(errs);
        return st;
      }
  }
  free(status);
  free(errs);
  return FHO_OK;
}

int fho_solve(const double *nodes, size_t n_nodes, const fho_segment *segments,
              size_t n_segments, const fho_grid *grids, const fho_port *ports, size_t n_ports,
              const double *freqs, size_t K, double *z_out, int threads) {
  fho_system *s;
  int st = fho_system_create(nodes, n_nodes, segments, n_segments, grids, ports, n_ports,
                             threads, &s);
  if (st) return st;
  st = fho_system_sweep(s, freqs, K, z_out, threads);
  fho_system_destroy(s);
  return st;
}

/* ======================================================================== */
/* pfft/{mod,grid,near,fft}.rs: the precorrected-FFT operator                */
/* ======================================================================== */

#define PFFT_QUAD_TOL 1e-10       /* grid.rs:41 */
#define PFFT_QUAD_MAX 12          /* grid.rs:43 */
#define PFFT_MAX_FFT_POINTS ((size_t)1 << 26) /* mod.rs:150 */

struct fho_pfft {
  size_t n;
  int has_grid;
  double origin[3], h;
  size_t dims[3], padded[3];
  int m;
  double *dirs; /* n x 3 */
  size_t *proj_off;
  uint32_t *proj_idx;
  double *proj_w;
  size_t *spread_off;
  uint32_t *spread_fil;
  double *spread_w;
  int comps[3], ncomp;
  size_t *near_off;
  uint32_t *near_col;
  double *near_val;
  double *spectrum; /* padded, natural order, real, includes 1/M */
};

/* mod.rs:614-631 */
double fho_auto_spacing(const double extent[3], double cells) {
  const double largest = fmax(fmax(extent[0], extent[1]), extent[2]);
  if (cells <= 1.0 || largest <= 0.0) return fmax(largest, 2.2250738585072014e-308);
  double small = largest / cells, large = largest;
  for (int it = 0; it < 200; ++it) {
    const double mid = sqrt(small * large);
    double count = 1.0;
    for (int a = 0; a < 3; ++a) count *= fmax(extent[a] / mid, 1.0);
    if (count > cells) small = mid;
    else large = mid;
  }
  return large;
}

/* fft.rs:59-80 */
size_t fho_smooth_size(size_t n) {
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

/* grid.rs:95-116: first plane and the m Lagrange weights. */
static size_t axis_stencil(const double origin[3], double h, const size_t dims[3], int m, int axis,
                           double x, double *weights) {
  const double u = (x - origin[axis]) / h;
  double fl = floor(u - 0.5 * (double)m) + 1.0;
  if (fl < 0.0) fl = 0.0;
  size_t first = (size_t)fl;
  if (first > dims[axis] - (size_t)m) first = dims[axis] - (size_t)m;
  const double t = u - (double)first;
  for (int k = 0; k < m; ++k) {
    double value = 1.0;
    for (int l = 0; l < m; ++l)
      if (l != k) value *= (t - (double)l) / ((double)k - (double)l);
    weights[k] = value;
  }
  return first;
}

size_t fho_axis_stencil(const double origin[3], double h, const size_t dims[3], int m, int axis,
                        double x, double *weights) {
  return axis_stencil(origin, h, dims, m, axis, x, weights);
}

typedef struct {
  size_t count;
  uint32_t *idx;
  double *w;
  size_t lo[3], hi[3];
} projection;

/* grid.rs:155-216.  Contributions are accumulated in node order (the
 * reference merges them after an unstable sort of the indices, so its order
 * of addition for a grid point is unspecified); entries are the grid points
 * some node's stencil touches, in ascending index. */
static int project_one(const struct fho_pfft *op, const fho_filament *f, double far,
                       const double *thresholds, projection *out) {
  const v3 axes[3] = {vload(f->dir), vload(f->wdir), vload(f->hdir)};
  const double extents[3] = {f->length, f->width, f->height};
  const v3 corner = vsub(vsub(vload(f->start), vmul(axes[1], 0.5 * f->width)),
                         vmul(axes[2], 0.5 * f->height));
  double *nx[3], *nw[3];
  size_t cnt[3];
  for (int d = 0; d < 3; ++d) {
    double pc = ceil(extents[d] / op->h);
    if (pc < 1.0) pc = 1.0;
    const size_t pieces = (size_t)pc;
    const double piece = extents[d] / (double)pieces;
    int order = order_for(thresholds, PFFT_QUAD_MAX, 0.5 * piece, far);
    if (!order) order = PFFT_QUAD_MAX;
    cnt[d] = pieces * (size_t)order;
    nx[d] = malloc(sizeof(double) * cnt[d]);
    nw[d] = malloc(sizeof(double) * cnt[d]);
    size_t q = 0;
    for (size_t p = 0; p < pieces; ++p)
      for (int i = 0; i < order; ++i) {
        rule_on(order, i, (double)p * piece, (double)(p + 1) * piece, &nx[d][q], &nw[d][q]);
        ++q;
      }
  }
  const double inverse_area = 1.0 / area_of(f);
  const int m = op->m;
  double wx[16], wy[16], wz[16];
  /* The union of the stencils: a box, found in a first pass. */
  size_t lo[3] = {SIZE_MAX, SIZE_MAX, SIZE_MAX}, hi[3] = {0, 0, 0};
  for (size_t a = 0; a < cnt[0]; ++a)
    for (size_t b = 0; b < cnt[1]; ++b)
      for (size_t c = 0; c < cnt[2]; ++c) {
        const v3 r = vadd(vadd(vadd(corner, vmul(axes[0], nx[0][a])), vmul(axes[1], nx[1][b])),
                          vmul(axes[2], nx[2][c]));
        const double rr[3] = {r.x, r.y, r.z};
        for (int ax = 0; ax < 3; ++ax) {
          const size_t first = axis_stencil(op->origin, op->h, op->dims, m, ax, rr[ax], wx);
          if (first < lo[ax]) lo[ax] = first;
          if (first + (size_t)m - 1 > hi[ax]) hi[ax] = first + (size_t)m - 1;
        }
      }
  const size_t bd[3] = {hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1};
  const size_t blen = bd[0] * bd[1] * bd[2];
  double *acc = calloc(blen, sizeof(double));
  unsigned char *touched = calloc(blen, 1);
  for (size_t a = 0; a < cnt[0]; ++a)
    for (size_t b = 0; b < cnt[1]; ++b)
      for (size_t c = 0; c < cnt[2]; ++c) {
        const v3 r = vadd(vadd(vadd(corner, vmul(axes[0], nx[0][a])), vmul(axes[1], nx[1][b])),
                          vmul(axes[2], nx[2][c]));
        const double weight = nw[0][a] * nw[1][b] * nw[2][c] * inverse_area;
        const size_t fx = axis_stencil(op->origin, op->h, op->dims, m, 0, r.x, wx);
        const size_t fy = axis_stencil(op->origin, op->h, op->dims, m, 1, r.y, wy);
        const size_t fz = axis_stencil(op->origin, op->h, op->dims, m, 2, r.z, wz);
        for (int i = 0; i < m; ++i)
          for (int j = 0; j < m; ++j) {
            const double lxy = weight * wx[i] * wy[j];
            for (int k = 0; k < m; ++k) {
              const size_t cell =
                  ((fx + i - lo[0]) * bd[1] + (fy + j - lo[1])) * bd[2] + (fz + k - lo[2]);
              acc[cell] += lxy * wz[k];
              touched[cell] = 1;
            }
          }
      }
  size_t count = 0;
  for (size_t q = 0; q < blen; ++q) count += touched[q];
  out->count = count;
  out->idx = malloc(sizeof(uint32_t) * (count ? count : 1));
  out->w = malloc(sizeof(double) * (count ? count : 1));
  size_t e = 0;
  for (size_t x = 0; x < bd[0]; ++x)
    for (size_t y = 0; y < bd[1]; ++y)
      for (size_t z = 0; z < bd[2]; ++z) {
        const size_t cell = (x * bd[1] + y) * bd[2] + z;
        if (!touched[cell]) continue;
        out->idx[e] =
            (uint32_t)(((lo[0] + x) * op->dims[1] + (lo[1] + y)) * op->dims[2] + (lo[2] + z));
        out->w[e] = acc[cell];
        ++e;
      }
  /* bounding box of the entries (grid.rs:202-214) */
  for (int ax = 0; ax < 3; ++ax) {
    out->lo[ax] = SIZE_MAX;
    out->hi[ax] = 0;
  }
  for (size_t q = 0; q < count; ++q) {
    const size_t g = out->idx[q];
    const size_t pz = g % op->dims[2], rest = g / op->dims[2];
    const size_t p[3] = {rest / op->dims[1], rest % op->dims[1], pz};
    for (int ax = 0; ax < 3; ++ax) {
      if (p[ax] < out->lo[ax]) out->lo[ax] = p[ax];
      if (p[ax] > out->hi[ax]) out->hi[ax] = p[ax];
    }
  }
  free(acc);
  free(touched);
  for (int d = 0; d < 3; ++d) {
    free(nx[d]);
    free(nw[d]);
  }
  return FHO_OK;
}

/* neumann.rs bounds (near.rs:27-34) */
static void bar_bounds(const bar *b, v3 *lo, v3 *hi) {
  v3 reach = V(0, 0, 0);
  for (int k = 0; k < 3; ++k) reach = vadd(reach, vmul(vabs(b->axes[k]), b->half[k]));
  *lo = vsub(b->centre, reach);
  *hi = vadd(b->centre, reach);
}

static int cmp_u32(const void *a, const void *b) {
  const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  return x < y ? -1 : x > y;
}

/* near.rs:38-109: for each i, the j >= i within radius (and i), ascending. */
static void near_pairs(const fho_filament *f, size_t n, double radius, size_t **off_out,
                       uint32_t **idx_out) {
  bar *bars = malloc(sizeof(bar) * n);
  v3 *blo = malloc(sizeof(v3) * n), *bhi = malloc(sizeof(v3) * n);
  for (size_t i = 0; i < n; ++i) {
    bars[i] = bar_new(&f[i]);
    bar_bounds(&bars[i], &blo[i], &bhi[i]);
  }
  v3 lo = blo[0], hi = bhi[0];
  for (size_t i = 0; i < n; ++i) {
    lo = V(fmin(lo.x, blo[i].x), fmin(lo.y, blo[i].y), fmin(lo.z, blo[i].z));
    hi = V(fmax(hi.x, bhi[i].x), fmax(hi.y, bhi[i].y), fmax(hi.z, bhi[i].z));
  }
  const v3 extent = vsub(hi, lo);
  const double e3[3] = {extent.x, extent.y, extent.z};
  const double emax = fmax(fmax(extent.x, extent.y), extent.z);
  double size = fmax(fmax(radius, emax * 1e-9), 2.2250738585072014e-308);
  size_t dims[3];
  for (;;) {
    size_t prod = 1;
    for (int a = 0; a < 3; ++a) {
      dims[a] = (size_t)floor(e3[a] / size) + 1;
      prod *= dims[a];
    }
    if (prod <= 8 * n + 64) break;
    size *= 1.5;
  }
  const double l3[3] = {lo.x, lo.y, lo.z};
#define CELL(x, a) \
  ((size_t)fmin(fmax(floor(((x) - l3[a]) / size), 0.0), (double)(dims[a] - 1)))
  const size_t nb = dims[0] * dims[1] * dims[2];
  /* bins: counting pass, then filling in filament order */
  size_t *bstart = calloc(nb + 1, sizeof(size_t));
  for (size_t i = 0; i < n; ++i) {
    const size_t f0[3] = {CELL(blo[i].x, 0), CELL(blo[i].y, 1), CELL(blo[i].z, 2)};
    const size_t f1[3] = {CELL(bhi[i].x, 0), CELL(bhi[i].y, 1), CELL(bhi[i].z, 2)};
    for (size_t a = f0[0]; a <= f1[0]; ++a)
      for (size_t b = f0[1]; b <= f1[1]; ++b)
        for (size_t c = f0[2]; c <= f1[2]; ++c) bstart[(a * dims[1] + b) * dims[2] + c + 1]++;
  }
  for (size_t q = 0; q < nb; ++q) bstart[q + 1] += bstart[q];
  size_t *fillp = malloc(sizeof(size_t) * (nb ? nb : 1));
  memcpy(fillp, bstart, sizeof(size_t) * nb);
  uint32_t *items = malloc(sizeof(uint32_t) * (bstart[nb] ? bstart[nb] : 1));
  for (size_t i = 0; i < n; ++i) {
    const size_t f0[3] = {CELL(blo[i].x, 0), CELL(blo[i].y, 1), CELL(blo[i].z, 2)};
    const size_t f1[3] = {CELL(bhi[i].x, 0), CELL(bhi[i].y, 1), CELL(bhi[i].z, 2)};
    for (size_t a = f0[0]; a <= f1[0]; ++a)
      for (size_t b = f0[1]; b <= f1[1]; ++b)
        for (size_t c = f0[2]; c <= f1[2]; ++c)
          items[fillp[(a * dims[1] + b) * dims[2] + c]++] = (uint32_t)i;
  }
  uint32_t **rows = malloc(sizeof(uint32_t *) * (n ? n : 1));
  size_t *rlen = malloc(sizeof(size_t) * (n ? n : 1));
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
    const size_t i = (size_t)ii;
    const v3 a0 = vsub(blo[i], V(radius, radius, radius));
    const v3 a1 = vadd(bhi[i], V(radius, radius, radius));
    const size_t f0[3] = {CELL(a0.x, 0), CELL(a0.y, 1), CELL(a0.z, 2)};
    const size_t f1[3] = {CELL(a1.x, 0), CELL(a1.y, 1), CELL(a1.z, 2)};
    size_t cap = 64, len = 0;
    uint32_t *found = malloc(sizeof(uint32_t) * cap);
    for (size_t a = f0[0]; a <= f1[0]; ++a)
      for (size_t b = f0[1]; b <= f1[1]; ++b)
        for (size_t c = f0[2]; c <= f1[2]; ++c) {
          const size_t q = (a * dims[1] + b) * dims[2] + c;
          for (size_t e = bstart[q]; e < bstart[q + 1]; ++e) {
            if (items[e] < i) continue;
            if (len == cap) {
              cap *= 2;
              found = realloc(found, sizeof(uint32_t) * cap);
            }
            found[len++] = items[e];
          }
        }
    qsort(found, len, sizeof(uint32_t), cmp_u32);
    size_t w = 0;
    for (size_t e = 0; e < len; ++e) {
      if (e > 0 && found[e] == found[e - 1]) continue; /* dedup on the sorted input */
      const uint32_t other = found[e];
      if (other == i || separation(&bars[i], &bars[other]) < radius) found[w++] = other;
    }
    rows[i] = found;
    rlen[i] = w;
  }
#undef CELL
  size_t *off = malloc(sizeof(size_t) * (n + 1));
  off[0] = 0;
  for (size_t i = 0; i < n; ++i) off[i + 1] = off[i] + rlen[i];
  uint32_t *idx = malloc(sizeof(uint32_t) * (off[n] ? off[n] : 1));
  for (size_t i = 0; i < n; ++i) {
    memcpy(idx + off[i], rows[i], sizeof(uint32_t) * rlen[i]);
    free(rows[i]);
  }
  free(rows);
  free(rlen);
  free(fillp);
  free(items);
  free(bstart);
  free(bars);
  free(blo);
  free(bhi);
  *off_out = off;
  *idx_out = idx;
}

/* near.rs:131-139 */
static double grid_kernel(int64_t dx, int64_t dy, int64_t dz, double h) {
  const double r2 = (double)(dx * dx + dy * dy + dz * dz);
  return r2 == 0.0 ? 0.0 : 1.0 / (h * sqrt(r2));
}

/* near.rs:316-389: sum_g sum_h W_ig W_jh K(g - h) for every partner j. */
static void grid_interactions(const struct fho_pfft *op, const projection *P, size_t i,
                              const size_t *partners, size_t np, const int64_t reach[3],
                              const size_t tdims[3], const double *table, double *out) {
  const projection *src = &P[i];
  size_t lo[3] = {SIZE_MAX, SIZE_MAX, SIZE_MAX}, hi[3] = {0, 0, 0};
  size_t partner_entries = 0;
  for (size_t q = 0; q < np; ++q) {
    const projection *pj = &P[partners[q]];
    for (int a = 0; a < 3; ++a) {
      if (pj->lo[a] < lo[a]) lo[a] = pj->lo[a];
      if (pj->hi[a] > hi[a]) hi[a] = pj->hi[a];
    }
    partner_entries += pj->count;
  }
  const size_t bd[3] = {hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1};
  const size_t box_len = bd[0] * bd[1] * bd[2];
  const size_t direct_cost = src->count * partner_entries;
  const size_t box_cost = src->count * box_len + partner_entries;
  const size_t d1 = op->dims[1], d2 = op->dims[2];
#define PT(g, p)                                   \
  do {                                             \
    const size_t z_ = (g) % d2, r_ = (g) / d2;     \
    (p)[0] = (int64_t)(r_ / d1);                   \
    (p)[1] = (int64_t)(r_ % d1);                   \
    (p)[2] = (int64_t)z_;                          \
  } while (0)
#define TIDX(d0, d1_, d2_) \
  ((size_t)(((d0) + reach[0]) * (int64_t)tdims[1] + ((d1_) + reach[1])) * tdims[2] + (size_t)((d2_) + reach[2]))
  if (direct_cost <= box_cost) {
    for (size_t q = 0; q < np; ++q) {
      const projection *pj = &P[partners[q]];
      double total = 0.0;
      for (size_t e = 0; e < pj->count; ++e) {
        int64_t hp[3];
        PT(pj->idx[e], hp);
        double potential = 0.0;
        for (size_t s = 0; s < src->count; ++s) {
          int64_t gp[3];
          PT(src->idx[s], gp);
          potential += src->w[s] * table[TIDX(hp[0] - gp[0], hp[1] - gp[1], hp[2] - gp[2])];
        }
        total += pj->w[e] * potential;
      }
      out[q] = total;
    }
    return;
  }
  double *psi = calloc(box_len, sizeof(double));
  const size_t bx = bd[0], by = bd[1], bz = bd[2];
  for (size_t s = 0; s < src->count; ++s) {
    int64_t g[3];
    PT(src->idx[s], g);
    const double wg = src->w[s];
    for (size_t x = 0; x < bx; ++x)
      for (size_t y = 0; y < by; ++y) {
        const double *kernel =
            &table[TIDX((int64_t)(lo[0] + x) - g[0], (int64_t)(lo[1] + y) - g[1], (int64_t)lo[2] - g[2])];
        double *row = &psi[(x * by + y) * bz];
        /* near.rs:189-207: fused multiply-add in blocks of four, plain
         * multiply-add for the tail */
        const size_t body = bz - bz % 4;
        for (size_t k = 0; k < body; ++k) row[k] = fma(wg, kernel[k], row[k]);
        for (size_t k = body; k < bz; ++k) row[k] += wg * kernel[k];
      }
  }
  for (size_t q = 0; q < np; ++q) {
    const projection *pj = &P[partners[q]];
    double total = 0.0;
    for (size_t e = 0; e < pj->count; ++e) {
      int64_t p[3];
      PT(pj->idx[e], p);
      const size_t local = (((size_t)p[0] - lo[0]) * by + ((size_t)p[1] - lo[1])) * bz + ((size_t)p[2] - lo[2]);
      total += pj->w[e] * psi[local];
    }
    out[q] = total;
  }
  free(psi);
#undef PT
#undef TIDX
}

/* ---- FFT (mixed radix 2, 3, 5; plain DFT for any other factor) ---- */
static void fft1(cplx *x, cplx *work, size_t n, size_t stride, int inverse, const cplx *tw,
                 size_t tw_n) {
  /* x: n values at stride; result written back in place (stride). */
  if (n == 1) return;
  size_t p = 2;
  if (n % 2) p = n % 3 == 0 ? 3 : (n % 5 == 0 ? 5 : n);
  const size_t q = n / p;
  /* gather the p decimated sub-sequences contiguously into work, transform each */
  for (size_t r = 0; r < p; ++r)
    for (size_t k = 0; k < q; ++k) work[r * q + k] = x[(k * p + r) * stride];
  for (size_t r = 0; r < p; ++r) fft1(work + r * q, work + n, q, 1, inverse, tw, tw_n);
  const size_t step = tw_n / n;
  for (size_t k = 0; k < q; ++k)
    for (size_t s = 0; s < p; ++s) {
      const size_t out = k + s * q;
      cplx acc = {0.0, 0.0};
      for (size_t r = 0; r < p; ++r) {
        const size_t e = (r * out) % n;
        cplx w = tw[e * step];
        if (inverse) w.im = -w.im;
        const cplx v = cmul(w, work[r * q + k]);
        acc.re += v.re;
        acc.im += v.im;
      }
      x[out * stride] = acc;
    }
}

/* in-place 1-D transform of every line along `axis` of a d0 x d1 x d2 array */
static void fft_axis(cplx *data, const size_t d[3], int axis, int inverse) {
  const size_t n = d[axis];
  if (n <= 1) return;
  cplx *tw = malloc(sizeof(cplx) * n);
  for (size_t k = 0; k < n; ++k) {
    const double a = -2.0 * M_PI * (double)k / (double)n;
    tw[k].re = cos(a);
    tw[k].im = sin(a);
  }
  const size_t stride = axis == 2 ? 1 : (axis == 1 ? d[2] : d[1] * d[2]);
  const size_t lines = d[0] * d[1] * d[2] / n;
#pragma omp parallel
  {
    cplx *line = malloc(sizeof(cplx) * n);
    cplx *work = malloc(sizeof(cplx) * 4 * n);
#pragma omp for schedule(static)
    for (int64_t li = 0; li < (int64_t)lines; ++li) {
      size_t base;
      if (axis == 2) base = (size_t)li * n;
      else if (axis == 1) base = ((size_t)li / d[2]) * d[1] * d[2] + (size_t)li % d[2];
      else base = (size_t)li;
      for (size_t k = 0; k < n; ++k) line[k] = data[base + k * stride];
      fft1(line, work, n, 1, inverse, tw, n);
      for (size_t k = 0; k < n; ++k) data[base + k * stride] = line[k];
    }
    free(line);
    free(work);
  }
  free(tw);
}

static void fft3(cplx *data, const size_t d[3], int inverse) {
  if (!inverse) {
    fft_axis(data, d, 2, 0);
    fft_axis(data, d, 1, 0);
    fft_axis(data, d, 0, 0);
  } else {
    fft_axis(data, d, 0, 1);
    fft_axis(data, d, 1, 1);
    fft_axis(data, d, 2, 1);
  }
}

/* fft.rs:91-140: the kernel's spectrum on the padded grid, wrapped offsets. */
static void build_spectrum(struct fho_pfft *op) {
  const size_t *N = op->dims, *M = op->padded;
  const size_t total = M[0] * M[1] * M[2];
  cplx *data = calloc(total, sizeof(cplx));
  for (size_t i = 0; i < M[0]; ++i) {
    int64_t di;
    if (i < N[0]) di = (int64_t)i;
    else if (i + N[0] > M[0]) di = (int64_t)i - (int64_t)M[0];
    else continue;
    for (size_t j = 0; j < M[1]; ++j) {
      int64_t dj;
      if (j < N[1]) dj = (int64_t)j;
      else if (j + N[1] > M[1]) dj = (int64_t)j - (int64_t)M[1];
      else continue;
      for (size_t k = 0; k < M[2]; ++k) {
        int64_t dk;
        if (k < N[2]) dk = (int64_t)k;
        else if (k + N[2] > M[2]) dk = (int64_t)k - (int64_t)M[2];
        else continue;
        data[(i * M[1] + j) * M[2] + k].re = grid_kernel(di, dj, dk, op->h);
      }
    }
  }
  fft3(data, M, 0);
  const double scale = 1.0 / (double)total;
  op->spectrum = malloc(sizeof(double) * total);
  for (size_t q = 0; q < total; ++q) op->spectrum[q] = data[q].re * scale;
  free(data);
}

/* fft.rs:155-220: phi_a, phi_b = K * a, K * b (b may be NULL). */
static void convolve(const struct fho_pfft *op, const double *a, const double *b, double *pa,
                     double *pb) {
  const size_t *N = op->dims, *M = op->padded;
  const size_t total = M[0] * M[1] * M[2];
  cplx *data = calloc(total, sizeof(cplx));
  for (size_t i = 0; i < N[0]; ++i)
    for (size_t j = 0; j < N[1]; ++j)
      for (size_t k = 0; k < N[2]; ++k) {
        const size_t s = (i * N[1] + j) * N[2] + k;
        cplx *d = &data[(i * M[1] + j) * M[2] + k];
        d->re = a[s];
        d->im = b ? b[s] : 0.0;
      }
  fft3(data, M, 0);
  for (size_t q = 0; q < total; ++q) {
    data[q].re *= op->spectrum[q];
    data[q].im *= op->spectrum[q];
  }
  fft3(data, M, 1);
  for (size_t i = 0; i < N[0]; ++i)
    for (size_t j = 0; j < N[1]; ++j)
      for (size_t k = 0; k < N[2]; ++k) {
        const size_t s = (i * N[1] + j) * N[2] + k;
        const cplx d = data[(i * M[1] + j) * M[2] + k];
        pa[s] = d.re;
        if (pb) pb[s] = d.im;
      }
  free(data);
}

int fho_pfft_convolve(const fho_pfft *op, const double *a, const double *b, double *pa,
                      double *pb) {
  if (!op->has_grid) return FHO_E_INVALID_ARGUMENT;
  convolve(op, a, b, pa, pb);
  return FHO_OK;
}

void fho_pfft_destroy(fho_pfft *op) {
  if (!op) return;
  free(op->dirs);
  free(op->proj_off);
  free(op->proj_idx);
  free(op->proj_w);
  free(op->spread_off);
  free(op->spread_fil);
  free(op->spread_w);
  free(op->near_off);
  free(op->near_col);
  free(op->near_val);
  free(op->spectrum);
  free(op);
}

/* mod.rs:332-445 */
int fho_pfft_create(const fho_filament *f, size_t n, const fho_pfft_params *prm, fho_pfft **out) {
  ensure_tables();
  *out = NULL;
  const int order = prm->interpolation_order;
  if (order < 1) return fail(FHO_E_INVALID_ARGUMENT, 1, 0, 0, 0.0);
  if (!(isfinite(prm->near_field_radius) && prm->near_field_radius >= (double)(order + 1)))
    return fail(FHO_E_INVALID_ARGUMENT, 2, 0, 0, 0.0);
  if (prm->auto_grid) {
    if (!(prm->cells_per_filament > 0.0 && isfinite(prm->cells_per_filament)))
      return fail(FHO_E_INVALID_ARGUMENT, 3, 0, 0, 0.0);
  } else if (!(prm->spacing > 0.0 && isfinite(prm->spacing))) {
    return fail(FHO_E_INVALID_ARGUMENT, 4, 0, 0, 0.0);
  }
  struct fho_pfft *op = calloc(1, sizeof(struct fho_pfft));
  op->n = n;
  op->m = order + 1;
  op->proj_off = calloc(n + 1, sizeof(size_t));
  op->near_off = calloc(n + 1, sizeof(size_t));
  op->dirs = malloc(sizeof(double) * 3 * (n ? n : 1));
  if (n == 0) {
    *out = op;
    return FHO_OK;
  }
  /* mod.rs:580-593 bounding box */
  v3 lo = V(INFINITY, INFINITY, INFINITY), hi = V(-INFINITY, -INFINITY, -INFINITY);
  for (size_t i = 0; i < n; ++i) {
    const v3 reach =
        vmul(vadd(vmul(vabs(vload(f[i].wdir)), f[i].width), vmul(vabs(vload(f[i].hdir)), f[i].height)), 0.5);
    const v3 ends[2] = {vload(f[i].start), vload(f[i].end)};
    for (int e = 0; e < 2; ++e) {
      const v3 a = vsub(ends[e], reach), b = vadd(ends[e], reach);
      lo = V(fmin(lo.x, a.x), fmin(lo.y, a.y), fmin(lo.z, a.z));
      hi = V(fmax(hi.x, b.x), fmax(hi.y, b.y), fmax(hi.z, b.z));
    }
  }
  const double ext[3] = {hi.x - lo.x, hi.y - lo.y, hi.z - lo.z};
  const double h = prm->auto_grid ? fho_auto_spacing(ext, prm->cells_per_filament * (double)n)
                                  : prm->spacing;
  /* grid.rs:64-80 */
  const double pad = (double)((op->m + 1) / 2);
  const double lo3[3] = {lo.x, lo.y, lo.z}, hi3[3] = {hi.x, hi.y, hi.z};
  size_t padded_total = 1, grid_total = 1;
  for (int a = 0; a < 3; ++a) {
    op->origin[a] = lo3[a] - pad * h;
    const double span = (hi3[a] - op->origin[a]) / h;
    op->dims[a] = (size_t)floor(span + 0.5 * (double)op->m) + 2;
    op->padded[a] = fho_smooth_size(2 * op->dims[a] - 1);
    padded_total *= op->padded[a];
    grid_total *= op->dims[a];
  }
  op->h = h;
  if (padded_total > PFFT_MAX_FFT_POINTS || grid_total > UINT32_MAX) {
    fho_pfft_destroy(op);
    return fail(FHO_E_INVALID_ARGUMENT, 5, (int64_t)padded_total, (int64_t)PFFT_MAX_FFT_POINTS, 0.0);
  }
  op->has_grid = 1;
  const double radius = prm->near_field_radius * h;
  /* projections */
  double thresholds[PFFT_QUAD_MAX];
  table_new(PFFT_QUAD_TOL, PFFT_QUAD_MAX, thresholds);
  projection *P = calloc(n, sizeof(projection));
#pragma omp parallel for schedule(dynamic, 8)
  for (int64_t i = 0; i < (int64_t)n; ++i) project_one(op, &f[i], radius, thresholds, &P[i]);
  for (size_t i = 0; i < n; ++i) op->proj_off[i + 1] = op->proj_off[i] + P[i].count;
  const size_t E = op->proj_off[n];
  op->proj_idx = malloc(sizeof(uint32_t) * (E ? E : 1));
  op->proj_w = malloc(sizeof(double) * (E ? E : 1));
  for (size_t i = 0; i < n; ++i) {
    memcpy(op->proj_idx + op->proj_off[i], P[i].idx, sizeof(uint32_t) * P[i].count);
    memcpy(op->proj_w + op->proj_off[i], P[i].w, sizeof(double) * P[i].count);
  }
  /* transpose (mod.rs:395-417) */
  op->spread_off = calloc(grid_total + 1, sizeof(size_t));
  for (size_t e = 0; e < E; ++e) op->spread_off[op->proj_idx[e] + 1]++;
  for (size_t g = 0; g < grid_total; ++g) op->spread_off[g + 1] += op->spread_off[g];
  size_t *fillp = malloc(sizeof(size_t) * grid_total);
  memcpy(fillp, op->spread_off, sizeof(size_t) * grid_total);
  op->spread_fil = malloc(sizeof(uint32_t) * (E ? E : 1));
  op->spread_w = malloc(sizeof(double) * (E ? E : 1));
  for (size_t i = 0; i < n; ++i)
    for (size_t e = op->proj_off[i]; e < op->proj_off[i + 1]; ++e) {
      const size_t slot = fillp[op->proj_idx[e]]++;
      op->spread_fil[slot] = (uint32_t)i;
      op->spread_w[slot] = op->proj_w[e];
    }
  free(fillp);
  for (size_t i = 0; i < n; ++i) {
    op->dirs[3 * i] = f[i].dir[0];
    op->dirs[3 * i + 1] = f[i].dir[1];
    op->dirs[3 * i + 2] = f[i].dir[2];
  }
  op->ncomp = 0;
  for (int c = 0; c < 3; ++c) {
    int any = 0;
    for (size_t i = 0; i < n && !any; ++i) any = op->dirs[3 * i + c] != 0.0;
    if (any) op->comps[op->ncomp++] = c;
  }
  /* near pairs and precorrection (near.rs:213-313) */
  size_t *noff;
  uint32_t *nidx;
  near_pairs(f, n, radius, &noff, &nidx);
  size_t **partners = malloc(sizeof(size_t *) * n);
  size_t *np = calloc(n, sizeof(size_t));
  for (size_t i = 0; i < n; ++i) {
    partners[i] = malloc(sizeof(size_t) * (noff[i + 1] - noff[i] + 1));
    const v3 di = vload(f[i].dir);
    for (size_t e = noff[i]; e < noff[i + 1]; ++e) {
      const size_t j = nidx[e];
      if (fabs(vdot(di, vload(f[j].dir))) >= 1e-14) partners[i][np[i]++] = j;
    }
  }
  int64_t reach[3] = {0, 0, 0};
  for (size_t i = 0; i < n; ++i) {
    if (!np[i]) continue;
    size_t plo[3] = {SIZE_MAX, SIZE_MAX, SIZE_MAX}, phi_[3] = {0, 0, 0};
    for (size_t q = 0; q < np[i]; ++q)
      for (int a = 0; a < 3; ++a) {
        if (P[partners[i][q]].lo[a] < plo[a]) plo[a] = P[partners[i][q]].lo[a];
        if (P[partners[i][q]].hi[a] > phi_[a]) phi_[a] = P[partners[i][q]].hi[a];
      }
    for (int a = 0; a < 3; ++a) {
      const size_t own_lo = P[i].lo[a], own_hi = P[i].hi[a];
      const size_t r1 = phi_[a] - (own_lo < phi_[a] ? own_lo : phi_[a]);
      const size_t r2 = own_hi - (plo[a] < own_hi ? plo[a] : own_hi);
      const int64_t r = (int64_t)(r1 > r2 ? r1 : r2);
      if (r > reach[a]) reach[a] = r;
    }
  }
  const size_t tdims[3] = {2 * (size_t)reach[0] + 1, 2 * (size_t)reach[1] + 1, 2 * (size_t)reach[2] + 1};
  double *table = malloc(sizeof(double) * tdims[0] * tdims[1] * tdims[2]);
  for (size_t x = 0; x < tdims[0]; ++x)
    for (size_t y = 0; y < tdims[1]; ++y)
      for (size_t z = 0; z < tdims[2]; ++z)
        table[(x * tdims[1] + y) * tdims[2] + z] =
            grid_kernel((int64_t)x - reach[0], (int64_t)y - reach[1], (int64_t)z - reach[2], h);
  double **upper = malloc(sizeof(double *) * n);
  int status = FHO_OK;
  int64_t err_i = -1, err_j = -1;
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
    const size_t i = (size_t)ii;
    upper[i] = malloc(sizeof(double) * (np[i] ? np[i] : 1));
    if (!np[i]) continue;
    double *approx = malloc(sizeof(double) * np[i]);
    grid_interactions(op, P, i, partners[i], np[i], reach, tdims, table, approx);
    const v3 di = vload(f[i].dir);
    for (size_t q = 0; q < np[i]; ++q) {
      const size_t j = partners[i][q];
      double exact;
      int meth, res;
      if (fho_mutual(&f[i], &f[j], &exact, &meth, &res) != FHO_OK) {
#pragma omp critical
        {
          if (err_i < 0 || (int64_t)i < err_i || ((int64_t)i == err_i && (int64_t)j < err_j)) {
            err_i = (int64_t)i;
            err_j = (int64_t)j;
            status = FHO_E_KERNEL;
          }
        }
        exact = NAN;
      }
      const double cos_ = vdot(di, vload(f[j].dir));
      upper[i][q] = exact - 1.0e-7 * cos_ * approx[q];
    }
    free(approx);
  }
  free(table);
  if (status) {
    for (size_t i = 0; i < n; ++i) {
      free(upper[i]);
      free(partners[i]);
      free(P[i].idx);
      free(P[i].w);
    }
    free(upper);
    free(partners);
    free(np);
    free(P);
    free(noff);
    free(nidx);
    fho_pfft_destroy(op);
    return fail(FHO_E_KERNEL, 2, err_i, err_j, 0.0);
  }
  /* mirror into full rows (near.rs:264-300) */
  size_t *cnt = calloc(n, sizeof(size_t));
  for (size_t i = 0; i < n; ++i)
    for (size_t q = 0; q < np[i]; ++q) {
      cnt[i]++;
      if (partners[i][q] != i) cnt[partners[i][q]]++;
    }
  for (size_t i = 0; i < n; ++i) op->near_off[i + 1] = op->near_off[i] + cnt[i];
  const size_t NE = op->near_off[n];
  op->near_col = malloc(sizeof(uint32_t) * (NE ? NE : 1));
  op->near_val = malloc(sizeof(double) * (NE ? NE : 1));
  memcpy(cnt, op->near_off, sizeof(size_t) * n);
  for (size_t i = 0; i < n; ++i)
    for (size_t q = 0; q < np[i]; ++q) {
      const size_t j = partners[i][q];
      op->near_col[cnt[i]] = (uint32_t)j;
      op->near_val[cnt[i]++] = upper[i][q];
      if (j != i) {
        op->near_col[cnt[j]] = (uint32_t)i;
        op->near_val[cnt[j]++] = upper[i][q];
      }
    }
  free(cnt);
  for (size_t i = 0; i < n; ++i) {
    free(upper[i]);
    free(partners[i]);
    free(P[i].idx);
    free(P[i].w);
  }
  free(upper);
  free(partners);
  free(np);
  free(P);
  free(noff);
  free(nidx);
  build_spectrum(op);
  *out = op;
  return FHO_OK;
}

/* mod.rs:482-528 */
void fho_pfft_apply(const fho_pfft *op, const double *x, double *y) {
  const size_t n = op->n;
#pragma omp parallel for schedule(static)
  for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
    double s = 0.0;
    for (size_t e = op->near_off[ii]; e < op->near_off[ii + 1]; ++e)
      s += op->near_val[e] * x[op->near_col[e]];
    y[ii] = s;
  }
  if (!op->has_grid) return;
  const size_t G = op->dims[0] * op->dims[1] * op->dims[2];
  double *src[3] = {NULL, NULL, NULL}, *pot[3] = {NULL, NULL, NULL};
  for (int q = 0; q < op->ncomp; ++q) {
    const int c = op->comps[q];
    src[q] = malloc(sizeof(double) * G);
    pot[q] = malloc(sizeof(double) * G);
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < (int64_t)G; ++g) {
      double s = 0.0;
      for (size_t e = op->spread_off[g]; e < op->spread_off[g + 1]; ++e) {
        const size_t i = op->spread_fil[e];
        s += op->spread_w[e] * op->dirs[3 * i + c] * x[i];
      }
      src[q][g] = s;
    }
  }
  for (int q = 0; q < op->ncomp; q += 2)
    convolve(op, src[q], q + 1 < op->ncomp ? src[q + 1] : NULL, pot[q],
             q + 1 < op->ncomp ? pot[q + 1] : NULL);
#pragma omp parallel for schedule(static)
  for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
    double far = 0.0;
    for (int q = 0; q < op->ncomp; ++q) {
      double sum = 0.0;
      for (size_t e = op->proj_off[ii]; e < op->proj_off[ii + 1]; ++e)
        sum += op->proj_w[e] * pot[q][op->proj_idx[e]];
      far += op->dirs[3 * ii + op->comps[q]] * sum;
    }
    y[ii] += 1.0e-7 * far;
  }
  for (int q = 0; q < op->ncomp; ++q) {
    free(src[q]);
    free(pot[q]);
  }
}

void fho_pfft_stats(const fho_pfft *op, double *spacing, size_t dims[3], size_t padded[3],
                    size_t *projection_entries, size_t *near_entries) {
  *spacing = op->h;
  for (int a = 0; a < 3; ++a) {
    dims[a] = op->has_grid ? op->dims[a] : 0;
    padded[a] = op->has_grid ? op->padded[a] : 0;
  }
  *projection_entries = op->proj_off[op->n];
  *near_entries = op->near_off[op->n];
}

int fho_pfft_data(const fho_pfft *op, size_t *proj_off, uint32_t *proj_idx, double *proj_w,
                  size_t *near_off, uint32_t *near_col, double *near_val, double origin[3]) {
  if (proj_off) memcpy(proj_off, op->proj_off, sizeof(size_t) * (op->n + 1));
  if (proj_idx) memcpy(proj_idx, op->proj_idx, sizeof(uint32_t) * op->proj_off[op->n]);
  if (proj_w) memcpy(proj_w, op->proj_w, sizeof(double) * op->proj_off[op->n]);
  if (near_off) memcpy(near_off, op->near_off, sizeof(size_t) * (op->n + 1));
  if (near_col) memcpy(near_col, op->near_col, sizeof(uint32_t) * op->near_off[op->n]);
  if (near_val) memcpy(near_val, op->near_val, sizeof(double) * op->near_off[op->n]);
  if (origin)
    for (int a = 0; a < 3; ++a) origin[a] = op->origin[a];
  return FHO_OK;
}

/* ======================================================================== */
/* gmres.rs: restarted, right-preconditioned GMRES (MGS twice, Givens)      */
/* ======================================================================== */

/* Complex and real forms of gmres.rs:91-206 with its helpers (:209-257);
 * the real form is the same recurrence with every imaginary part zero. */
typedef struct {
  double re, im;
} gz;
static inline gz gz_add(gz a, gz b) { gz r = {a.re + b.re, a.im + b.im}; return r; }
static inline gz gz_sub(gz a, gz b) { gz r = {a.re - b.re, a.im - b.im}; return r; }
static inline gz gz_mul(gz a, gz b) {
  gz r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return r;
}
static inline gz gz_conj(gz a) { gz r = {a.re, -a.im}; return r; }
static inline gz gz_scale(gz a, double s) { gz r = {a.re * s, a.im * s}; return r; }
static inline gz gz_unscale(gz a, double s) { gz r = {a.re / s, a.im / s}; return r; }
static inline double gz_mod(gz a) { return hypot(a.re, a.im); }
static inline gz gz_div(gz a, gz b) { /* num-complex Div */
  const double n = b.re * b.re + b.im * b.im;
  gz r = {(a.re * b.re + a.im * b.im) / n, (a.im * b.re - a.re * b.im) / n};
  return r;
}
static inline gz gz_neg(gz a) { gz r = {-a.re, -a.im}; return r; }

static double gz_norm(const gz *v, size_t n, int complex_) {
  double scale = 0.0;
  for (size_t i = 0; i < n; ++i) scale = fmax(scale, complex_ ? gz_mod(v[i]) : fabs(v[i].re));
  if (scale == 0.0 || !isfinite(scale)) return scale;
  double sum = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const gz u = gz_unscale(v[i], scale);
    sum += complex_ ? u.re * u.re + u.im * u.im : u.re * u.re;
  }
  return scale * sqrt(sum);
}

static gz gz_dot(const gz *u, const gz *v, size_t n) {
  gz s = {0.0, 0.0};
  for (size_t i = 0; i < n; ++i) s = gz_add(s, gz_mul(gz_conj(u[i]), v[i]));
  return s;
}

static void gz_axpy(gz a, const gz *x, gz *y, size_t n) {
  for (size_t i = 0; i < n; ++i) y[i] = gz_add(y[i], gz_mul(a, x[i]));
}

/* gmres.rs:246-257 */
static void givens(gz a, gz b, int complex_, double *c, gz *s, gz *r) {
  const double abs_a = complex_ ? gz_mod(a) : fabs(a.re), abs_b = complex_ ? gz_mod(b) : fabs(b.re);
  if (abs_b == 0.0) {
    *c = 1.0;
    s->re = s->im = 0.0;
    *r = a;
    return;
  }
  if (abs_a == 0.0) {
    *c = 0.0;
    s->re = 1.0;
    s->im = 0.0;
    *r = b;
    return;
  }
  const double length = hypot(abs_a, abs_b);
  const gz phase = gz_unscale(a, abs_a);
  *c = abs_a / length;
  *s = gz_mul(phase, gz_unscale(gz_conj(b), length));
  *r = gz_scale(phase, length);
}

typedef void (*gz_op)(void *ctx, const gz *in, gz *out);

typedef struct {
  size_t iterations;
  double relative_residual;
  int converged;
} gz_outcome;

static void gz_residual(gz_op apply, void *actx, const gz *b, const gz *x, gz *r, size_t n) {
  apply(actx, x, r);
  for (size_t i = 0; i < n; ++i) r[i] = gz_sub(b[i], r[i]);
}

static gz_outcome gmres_core(gz_op apply, void *actx, gz_op precondition, void *pctx,
                             const gz *b, gz *x, size_t n, double tolerance, size_t restart_in,
                             size_t max_iterations, int complex_) {
  gz_outcome out = {0, 0.0, 1};
  const double b_norm = gz_norm(b, n, complex_);
  if (b_norm == 0.0) {
    for (size_t i = 0; i < n; ++i) x[i].re = x[i].im = 0.0;
    return out;
  }
  const double target = tolerance * b_norm;
  size_t restart = restart_in < 1 ? 1 : restart_in;
  if (restart > n) restart = n;
  gz *residual = malloc(sizeof(gz) * n);
  gz *basis = malloc(sizeof(gz) * n * (restart + 1));
  gz *w = malloc(sizeof(gz) * n), *tmp = malloc(sizeof(gz) * n), *update = malloc(sizeof(gz) * n);
  gz *hess = malloc(sizeof(gz) * (restart + 1) * restart); /* column j: hess[j*(restart+1) + i] */
  double *rc = malloc(sizeof(double) * restart);
  gz *rs = malloc(sizeof(gz) * restart), *g = malloc(sizeof(gz) * (restart + 1));
  gz *y = malloc(sizeof(gz) * restart);
  size_t iterations = 0;
  gz_residual(apply, actx, b, x, residual, n);
  double r_norm = gz_norm(residual, n, complex_);
  for (;;) {
    if (r_norm <= target) {
      out.converged = 1;
      break;
    }
    if (iterations >= max_iterations || !isfinite(r_norm)) {
      out.converged = 0;
      break;
    }
    for (size_t i = 0; i < n; ++i) basis[i] = gz_unscale(residual[i], r_norm);
    size_t steps = 0, nb = 1;
    for (size_t i = 0; i <= restart; ++i) g[i].re = g[i].im = 0.0;
    g[0].re = r_norm;
    while (steps < restart && iterations < max_iterations) {
      const size_t j = steps;
      precondition(pctx, basis + j * n, tmp);
      apply(actx, tmp, w);
      ++iterations;
      gz *column = hess + j * (restart + 1);
      for (size_t i = 0; i < j + 2; ++i) column[i].re = column[i].im = 0.0;
      for (int pass = 0; pass < 2; ++pass)
        for (size_t v = 0; v < nb; ++v) {
          const gz projection = gz_dot(basis + v * n, w, n);
          column[v] = gz_add(column[v], projection);
          gz_axpy(gz_neg(projection), basis + v * n, w, n);
        }
      const double next_norm = gz_norm(w, n, complex_);
      column[j + 1].re = next_norm;
      column[j + 1].im = 0.0;
      for (size_t i = 0; i < j; ++i) {
        const gz a = column[i], bb = column[i + 1];
        column[i] = gz_add(gz_scale(a, rc[i]), gz_mul(rs[i], bb));
        column[i + 1] = gz_sub(gz_scale(bb, rc[i]), gz_mul(gz_conj(rs[i]), a));
      }
      double c;
      gz s, r;
      givens(column[j], column[j + 1], complex_, &c, &s, &r);
      column[j] = r;
      column[j + 1].re = column[j + 1].im = 0.0;
      rc[j] = c;
      rs[j] = s;
      const gz gj = g[j];
      g[j] = gz_scale(gj, c);
      g[j + 1] = gz_neg(gz_mul(gz_conj(s), gj));
      ++steps;
      if (next_norm == 0.0 || (complex_ ? gz_mod(g[j + 1]) : fabs(g[j + 1].re)) <= target) break;
      for (size_t i = 0; i < n; ++i) basis[nb * n + i] = gz_unscale(w[i], next_norm);
      ++nb;
    }
    for (size_t ii = steps; ii-- > 0;) {
      gz sum = g[ii];
      for (size_t k = ii + 1; k < steps; ++k)
        sum = gz_sub(sum, gz_mul(hess[k * (restart + 1) + ii], y[k]));
      y[ii] = gz_div(sum, hess[ii * (restart + 1) + ii]);
    }
    for (size_t i = 0; i < n; ++i) update[i].re = update[i].im = 0.0;
    for (size_t k = 0; k < steps; ++k) gz_axpy(y[k], basis + k * n, update, n);
    precondition(pctx, update, tmp);
    for (size_t i = 0; i < n; ++i) x[i] = gz_add(x[i], tmp[i]);
    gz_residual(apply, actx, b, x, residual, n);
    r_norm = gz_norm(residual, n, complex_);
  }
  out.iterations = iterations;
  out.relative_residual = r_norm / b_norm;
  free(residual);
  free(basis);
  free(w);
  free(tmp);
  free(update);
  free(hess);
  free(rc);
  free(rs);
  free(g);
  free(y);
  return out;
}

typedef struct {
  const gz *a;
  size_t n;
} dense_ctx;

static void dense_apply(void *ctx, const gz *in, gz *out) {
  const dense_ctx *d = ctx;
  for (size_t i = 0; i < d->n; ++i) {
    gz s = {0.0, 0.0};
    for (size_t k = 0; k < d->n; ++k) s = gz_add(s, gz_mul(d->a[k * d->n + i], in[k]));
    out[i] = s;
  }
}
static void identity_op(void *ctx, const gz *in, gz *out) {
  const size_t n = *(const size_t *)ctx;
  memcpy(out, in, sizeof(gz) * n);
}

/* GMRES on a dense column-major complex matrix (interleaved {re, im}), no
 * preconditioner: for checking gmres.rs against a direct solve. */
int fho_gmres_dense(size_t n, const double *a, const double *b, double *x, double tolerance,
                    size_t restart, size_t max_iterations, int complex_, size_t *iterations,
                    double *relative_residual, int *converged) {
  dense_ctx d = {(const gz *)a, n};
  size_t nn = n;
  const gz_outcome o = gmres_core(dense_apply, &d, identity_op, &nn, (const gz *)b, (gz *)x, n,
                                  tolerance, restart, max_iterations, complex_);
  *iterations = o.iterations;
  *relative_residual = o.relative_residual;
  *converged = o.converged;
  return FHO_OK;
}

/* ======================================================================== */
/* iterative.rs: matrix-free port impedance by GMRES on the pFFT operator   */
/* ======================================================================== */

#define EXACT_DIAGONAL_MAX_LOOP_LENGTH 32 /* iterative.rs:94 */

struct fho_iter {
  size_t counts[6];
  size_t n, m, p, loops;
  fho_filament *filaments;
  double *res;
  int64_t *row_ptr, *branch;
  double *sign;
  fho_pfft *op;
  double tolerance;
  size_t restart, max_iterations;
  double *r_ep, *l_ep, *r_pp, *l_pp; /* m x p, p x p row-major */
  double *r_diag, *l_diag;
  size_t last_iterations[64];
};

void fho_iter_destroy(fho_iter *s) {
  if (!s) return;
  free(s->filaments);
  free(s->res);
  free(s->row_ptr);
  free(s->branch);
  free(s->sign);
  fho_pfft_destroy(s->op);
  free(s->r_ep);
  free(s->l_ep);
  free(s->r_pp);
  free(s->l_pp);
  free(s->r_diag);
  free(s->l_diag);
  free(s);
}

/* iterative.rs:489-510: M_{0..loops} v */
static void gather_loops(const fho_iter *s, size_t loops, const double *v, double *out) {
  for (size_t i = 0; i < loops; ++i) {
    double sum = 0.0;
    for (int64_t e = s->row_ptr[i]; e < s->row_ptr[i + 1]; ++e) sum += v[s->branch[e]] * s->sign[e];
    out[i] = sum;
  }
}

/* iterative.rs:527-536 */
static int loop_inductance(const fho_iter *s, size_t row, double *out) {
  const int64_t b0 = s->row_ptr[row], b1 = s->row_ptr[row + 1];
  double own = 0.0;
  for (int64_t e = b0; e < b1; ++e) own += fho_self_inductance(&s->filaments[s->branch[e]]);
  if (b1 - b0 > EXACT_DIAGONAL_MAX_LOOP_LENGTH) {
    *out = own;
    return FHO_OK;
  }
  double mutual = 0.0;
  for (int64_t k = b0; k < b1; ++k)
    for (int64_t l = k + 1; l < b1; ++l) {
      double m;
      int meth, res;
      if (fho_mutual(&s->filaments[s->branch[k]], &s->filaments[s->branch[l]], &m, &meth, &res))
        return FHO_E_KERNEL;
      mutual += s->sign[k] * s->sign[l] * m;
    }
  *out = own + 2.0 * mutual;
  return FHO_OK;
}

/* iterative.rs:155-246 */
int fho_iter_create(const double *nodes, size_t n_nodes, const fho_segment *segments,
                    size_t n_segments, const fho_grid *grids, const fho_port *ports,
                    size_t n_ports, const fho_pfft_params *pp, double tolerance, size_t restart,
                    size_t max_iterations, fho_iter **out) {
  *out = NULL;
  size_t n;
  int st = fho_discretize_geometry(nodes, n_nodes, segments, n_segments, grids, NULL, 0, &n);
  if (st) return st;
  fho_iter *s = calloc(1, sizeof(fho_iter));
  s->filaments = malloc(sizeof(fho_filament) * (n ? n : 1));
  st = fho_discretize_geometry(nodes, n_nodes, segments, n_segments, grids, s->filaments, n, &n);
  if (st) goto fail_;
  size_t *per = malloc(sizeof(size_t) * (n_segments ? n_segments : 1));
  for (size_t q = 0; q < n_segments; ++q) per[q] = (size_t)grids[q].nw * grids[q].nh;
  size_t loops, internal, nnz;
  st = fho_mesh_build(n_nodes, segments, n_segments, per, ports, n_ports, NULL, NULL, NULL,
                      &loops, &internal, &nnz);
  if (st) {
    free(per);
    goto fail_;
  }
  s->row_ptr = malloc(sizeof(int64_t) * (loops + 1));
  s->branch = malloc(sizeof(int64_t) * (nnz ? nnz : 1));
  s->sign = malloc(sizeof(double) * (nnz ? nnz : 1));
  fho_mesh_build(n_nodes, segments, n_segments, per, ports, n_ports, s->row_ptr, s->branch,
                 s->sign, &loops, &internal, &nnz);
  free(per);
  s->n = n;
  s->m = internal;
  s->p = loops - internal;
  s->loops = loops;
  s->tolerance = tolerance;
  s->restart = restart;
  s->max_iterations = max_iterations;
  s->res = malloc(sizeof(double) * (n ? n : 1));
  for (size_t i = 0; i < n; ++i) {
    const fho_filament *f = &s->filaments[i];
    s->res[i] = f->length / (f->sigma * area_of(f));
  }
  st = fho_pfft_create(s->filaments, n, pp, &s->op);
  if (st) goto fail_;
  const size_t m = s->m, p = s->p;
  s->r_ep = calloc(m * p + 1, sizeof(double));
  s->l_ep = calloc(m * p + 1, sizeof(double));
  s->r_pp = calloc(p * p + 1, sizeof(double));
  s->l_pp = calloc(p * p + 1, sizeof(double));
  double *cur = malloc(sizeof(double) * (n ? n : 1)), *flux = malloc(sizeof(double) * (n ? n : 1));
  double *drop = malloc(sizeof(double) * (n ? n : 1));
  double *rc = malloc(sizeof(double) * loops), *lc = malloc(sizeof(double) * loops);
  for (size_t q = 0; q < p; ++q) {
    for (size_t b = 0; b < n; ++b) cur[b] = 0.0;
    for (int64_t e = s->row_ptr[m + q]; e < s->row_ptr[m + q + 1]; ++e) cur[s->branch[e]] += s->sign[e];
    fho_pfft_apply(s->op, cur, flux);
    for (size_t b = 0; b < n; ++b) drop[b] = cur[b] * s->res[b];
    gather_loops(s, loops, drop, rc);
    gather_loops(s, loops, flux, lc);
    for (size_t i = 0; i < m; ++i) {
      s->r_ep[i * p + q] = rc[i];
      s->l_ep[i * p + q] = lc[i];
    }
    for (size_t k = 0; k < p; ++k) {
      s->r_pp[k * p + q] = rc[m + k];
      s->l_pp[k * p + q] = lc[m + k];
    }
  }
  free(cur);
  free(flux);
  free(drop);
  free(rc);
  free(lc);
  /* (X + X^T) * 0.5 */
  double *t = malloc(sizeof(double) * (p * p + 1));
  for (int pass = 0; pass < 2; ++pass) {
    double *X = pass ? s->l_pp : s->r_pp;
    for (size_t a = 0; a < p; ++a)
      for (size_t b = 0; b < p; ++b) t[a * p + b] = (X[a * p + b] + X[b * p + a]) * 0.5;
    memcpy(X, t, sizeof(double) * p * p);
  }
  free(t);
  s->r_diag = malloc(sizeof(double) * (m ? m : 1));
  s->l_diag = malloc(sizeof(double) * (m ? m : 1));
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < (int64_t)m; ++i) {
    double r = 0.0;
    for (int64_t e = s->row_ptr[i]; e < s->row_ptr[i + 1]; ++e) r += s->res[s->branch[e]];
    s->r_diag[i] = r;
    if (loop_inductance(s, (size_t)i, &s->l_diag[i])) {
#pragma omp atomic write
      bad = 1;
    }
  }
  if (bad) {
    st = fail(FHO_E_KERNEL, 2, 0, 0, 0.0);
    goto fail_;
  }
  s->counts[0] = n_nodes;
  s->counts[1] = n_segments;
  s->counts[2] = n;
  s->counts[3] = loops;
  s->counts[4] = internal;
  s->counts[5] = loops - internal;
  *out = s;
  return FHO_OK;
fail_:
  fho_iter_destroy(s);
  return st;
}

typedef struct {
  const fho_iter *s;
  double omega;
  int dc;
  double *re, *im, *fre, *fim, *drop_re, *drop_im; /* branch scratch */
} ee_ctx;

/* iterative.rs:431-463 */
static void apply_ee(void *ctx, const gz *x, gz *y) {
  ee_ctx *c = ctx;
  const fho_iter *s = c->s;
  const size_t n = s->n, m = s->m;
  for (size_t b = 0; b < n; ++b) c->re[b] = c->im[b] = 0.0;
  for (size_t i = 0; i < m; ++i)
    for (int64_t e = s->row_ptr[i]; e < s->row_ptr[i + 1]; ++e) {
      c->re[s->branch[e]] += x[i].re * s->sign[e];
      c->im[s->branch[e]] += x[i].im * s->sign[e];
    }
  if (c->dc) {
    for (size_t b = 0; b < n; ++b) {
      c->drop_re[b] = c->re[b] * s->res[b];
      c->drop_im[b] = c->im[b] * s->res[b];
    }
  } else {
    fho_pfft_apply(s->op, c->re, c->fre);
    fho_pfft_apply(s->op, c->im, c->fim);
    for (size_t b = 0; b < n; ++b) {
      const double r = s->res[b];
      c->drop_re[b] = r * c->re[b] - c->omega * c->fim[b];
      c->drop_im[b] = r * c->im[b] + c->omega * c->fre[b];
    }
  }
  for (size_t i = 0; i < m; ++i) {
    gz sum = {0.0, 0.0};
    for (int64_t e = s->row_ptr[i]; e < s->row_ptr[i + 1]; ++e) {
      sum.re += c->drop_re[s->branch[e]] * s->sign[e];
      sum.im += c->drop_im[s->branch[e]] * s->sign[e];
    }
    y[i] = sum;
  }
}

typedef struct {
  const gz *d;
  size_t m;
} jacobi_ctx;
static void jacobi(void *ctx, const gz *v, gz *out) {
  const jacobi_ctx *j = ctx;
  for (size_t i = 0; i < j->m; ++i) out[i] = gz_div(v[i], j->d[i]);
}

/* iterative.rs:298-326 (solve) + 467-486 (schur_complement_gmres); z is
 * p x p complex interleaved.  iterations (p entries) may be NULL. */
int fho_iter_impedance(fho_iter *s, double frequency, double *z, size_t *iterations) {
  if (!(isfinite(frequency) && frequency >= 0.0))
    return fail(FHO_E_INVALID_FREQUENCY, 0, 0, 0, frequency);
  const size_t m = s->m, p = s->p, n = s->n;
  const int dc = frequency == 0.0;
  const double omega = 6.283185307179586 * frequency;
  gz *zz = malloc(sizeof(gz) * (p * p + 1));
  for (size_t a = 0; a < p; ++a)
    for (size_t b = 0; b < p; ++b) {
      zz[a * p + b].re = s->r_pp[a * p + b];
      zz[a * p + b].im = dc ? 0.0 : omega * s->l_pp[a * p + b];
    }
  if (m > 0) {
    gz *diag = malloc(sizeof(gz) * m);
    for (size_t i = 0; i < m; ++i) {
      diag[i].re = s->r_diag[i];
      diag[i].im = dc ? 0.0 : omega * s->l_diag[i];
    }
    ee_ctx c = {s, omega, dc, NULL, NULL, NULL, NULL, NULL, NULL};
    double *scratch = malloc(sizeof(double) * 6 * (n ? n : 1));
    c.re = scratch;
    c.im = scratch + n;
    c.fre = scratch + 2 * n;
    c.fim = scratch + 3 * n;
    c.drop_re = scratch + 4 * n;
    c.drop_im = scratch + 5 * n;
    jacobi_ctx jc = {diag, m};
    gz *rhs = malloc(sizeof(gz) * m), *x = malloc(sizeof(gz) * m);
    gz *solved = malloc(sizeof(gz) * m * p);
    for (size_t q = 0; q < p; ++q) {
      for (size_t i = 0; i < m; ++i) {
        rhs[i].re = s->r_ep[i * p + q];
        rhs[i].im = dc ? 0.0 : omega * s->l_ep[i * p + q];
        x[i].re = x[i].im = 0.0;
      }
      const gz_outcome o = gmres_core(apply_ee, &c, jacobi, &jc, rhs, x, m, s->tolerance,
                                      s->restart, s->max_iterations, !dc);
      if (iterations) iterations[q] = o.iterations;
      if (!o.converged) {
        free(diag);
        free(scratch);
        free(rhs);
        free(x);
        free(solved);
        free(zz);
        return fail(FHO_E_NOT_CONVERGED, 0, (int64_t)q, (int64_t)o.iterations, o.relative_residual);
      }
      for (size_t i = 0; i < m; ++i) solved[i * p + q] = x[i];
    }
    /* pp - ep^T solved */
    for (size_t a = 0; a < p; ++a)
      for (size_t b = 0; b < p; ++b) {
        gz acc = {0.0, 0.0};
        for (size_t i = 0; i < m; ++i) {
          const gz e = {s->r_ep[i * p + a], dc ? 0.0 : omega * s->l_ep[i * p + a]};
          acc = gz_add(acc, gz_mul(e, solved[i * p + b]));
        }
        zz[a * p + b] = gz_sub(zz[a * p + b], acc);
      }
    free(diag);
    free(scratch);
    free(rhs);
    free(x);
    free(solved);
  }
  for (size_t a = 0; a < p; ++a)
    for (size_t b = 0; b < p; ++b) {
      const gz e = gz_scale(gz_add(zz[a * p + b], zz[b * p + a]), 0.5);
      z[2 * (a * p + b)] = e.re;
      z[2 * (a * p + b) + 1] = dc ? 0.0 : e.im;
    }
  free(zz);
  for (size_t q = 0; q < 2 * p * p; ++q)
    if (!isfinite(z[q])) return fail(FHO_E_SINGULAR, 0, 0, 0, frequency);
  return FHO_OK;
}

void fho_iter_counts(const fho_iter *s, size_t counts[6]) { memcpy(counts, s->counts, sizeof(s->counts)); }
const fho_pfft *fho_iter_operator(const fho_iter *s) { return s->op; }
const fho_filament *fho_iter_filaments(const fho_iter *s) { return s->filaments; }
