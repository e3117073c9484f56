// Device log and atan built from IEEE basic operations only (+ - * /, which
// gfx950 rounds correctly in FP64; no fused multiply-add: every HIP source of
// this library is compiled with -ffp-contract=off).  The same algorithm on
// any IEEE-754 host gives the same bits, which is what lets the fill be
// checked bit for bit against a CPU evaluation.
//
//   log: x = 2^k m, m in [sqrt(1/2), sqrt(2)), f = m - 1 (exact),
//        s = f / (2 + f), log(1 + f) = 2 atanh(s) = f - f^2/2 + s (f^2/2 + R),
//        R = sum_{k=1..11} 2 s^(2k) / (2k + 1) (the atanh series), plus
//        k ln 2 as a 32-bit head (exact product) and a tail.
//   atan: reduction about c in {0, 1/2, 1, 3/2, inf} by
//        atan(x) = atan(c) + atan((x - c) / (1 + c x)), |u| <= 7/16, then the
//        Taylor series to u^49 (truncation below 1e-19); atan(c) as head +
//        tail from 113-bit arithmetic.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fh {

__device__ __forceinline__ double pm_log(double x) {
  if (!(x > 0.0)) {
    if (x == 0.0) return -__builtin_inf();
    return __builtin_nan("");
  }
  if (x == __builtin_inf()) return x;
  uint64_t b = __double_as_longlong(x);
  int k = 0;
  if (b < 0x0010000000000000ULL) {
    x *= 0x1p54;
    b = __double_as_longlong(x);
    k = -54;
  }
  k += static_cast<int>(b >> 52) - 1023;
  double m = __longlong_as_double(static_cast<long long>((b & 0x000fffffffffffffULL) |
                                                         0x3ff0000000000000ULL));
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
  const double dk = static_cast<double>(k);
  const double ln2_hi = 0x1.62e42feep-1, ln2_lo = 0x1.a39ef35793c76p-33;
  return dk * ln2_hi - ((hfsq - (s * (hfsq + r) + dk * ln2_lo)) - f);
}

__device__ __forceinline__ double pm_atan(double x) {
  if (x != x) return x;
  const double ax = __builtin_fabs(x);
  double hi, lo;
  int id;
  double u;
  if (ax >= 0x1p66) {
    const double r = 0x1.921fb54442d18p+0 + 0x1.1a62633145c07p-54;
    return x < 0 ? -r : r;
  }
  if (ax < 0.4375) {
    if (ax < 0x1p-27) return x;
    id = -1;
    u = ax;
    hi = lo = 0.0;
  } else if (ax < 0.6875) {
    id = 0;
    u = (2.0 * ax - 1.0) / (2.0 + ax);
    hi = 0x1.dac670561bb4fp-2;
    lo = 0x1.a2b7f222f65e2p-56;
  } else if (ax < 1.1875) {
    id = 1;
    u = (ax - 1.0) / (ax + 1.0);
    hi = 0x1.921fb54442d18p-1;
    lo = 0x1.1a62633145c07p-55;
  } else if (ax < 2.4375) {
    id = 2;
    u = (ax - 1.5) / (1.0 + 1.5 * ax);
    hi = 0x1.f730bd281f69bp-1;
    lo = 0x1.007887af0cbbdp-56;
  } else {
    id = 3;
    u = -1.0 / ax;
    hi = 0x1.921fb54442d18p+0;
    lo = 0x1.1a62633145c07p-54;
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
  p = z * p;
  double r;
  if (id < 0) {
    r = u + u * p;
  } else {
    r = hi - ((-(u * p) - lo) - u);
  }
  return x < 0 ? -r : r;
}

__device__ __forceinline__ double pm_hypot(double x, double y) {
  return __builtin_sqrt(x * x + y * y);
}

}  // namespace fh
