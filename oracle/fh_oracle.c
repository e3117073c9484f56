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
