// Double-precision natural logarithm and arctangent with a fixed operation
// sequence, so that a host build and a gfx950 build of the same expression
// give the same bits.  Both follow the published fdlibm method (Sun
// Microsystems, e_log.c and s_atan.c: reduction to a small interval, then an
// odd minimax polynomial with a split hi/lo constant), written here
// branch-free for wavefront execution: the reduction interval selects a
// numerator/denominator pair and one division serves every interval.
// Error: below 1 ulp on finite positive (log) / all finite (atan) inputs.
//
// Callers must compile these without FMA contraction (-ffp-contract=off) for
// the host/device results to agree; fma() is never used implicitly.
#pragma once
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DDD_HD __host__ __device__ __forceinline__
#else
#define DDD_HD inline
#endif

namespace ddd_libm {

DDD_HD uint64_t bits_of(double x) {
  uint64_t u;
  memcpy(&u, &x, sizeof(u));
  return u;
}
DDD_HD double from_bits(uint64_t u) {
  double x;
  memcpy(&x, &u, sizeof(x));
  return x;
}

// IEEE division; the device-only ranged forms substitute another division
// that gives the same bits on the operands these functions pass it.
struct IeeeDiv {
  DDD_HD double operator()(double a, double b) const { return a / b; }
};

// log(x) + k ln 2 for positive normal x: reduction to [sqrt(2)/2, sqrt(2)),
// s = f / (2 + f) (through div), the polynomial in s^2.
template <class Div> DDD_HD double log_core(double x, int k, Div div) {
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  const double Lg1 = 6.666666666666735130e-01, Lg2 = 3.999999999940941908e-01,
               Lg3 = 2.857142874366239149e-01, Lg4 = 2.222219843214978396e-01,
               Lg5 = 1.818357216161805012e-01, Lg6 = 1.531383769920937332e-01,
               Lg7 = 1.479819860511658591e-01;
  const uint64_t u = bits_of(x);
  // Normalise the mantissa into [sqrt(2)/2, sqrt(2)).
  uint32_t hx = (uint32_t)(u >> 32);
  hx += 0x3ff00000u - 0x3fe6a09eu;
  k += (int)(hx >> 20) - 0x3ff;
  hx = (hx & 0x000fffffu) + 0x3fe6a09eu;
  const double m = from_bits(((uint64_t)hx << 32) | (u & 0xffffffffull));
  const double f = m - 1.0;
  const double hfsq = 0.5 * f * f;
  const double s = div(f, 2.0 + f);
  const double z = s * s;
  const double w = z * z;
  const double t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
  const double t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
  const double R = t2 + t1;
  const double dk = (double)k;
  return s * (hfsq + R) + dk * ln2_lo - hfsq + f + dk * ln2_hi;
}

// log(x) for x > 0 finite; other inputs follow IEEE conventions
// (log(0) = -inf, log(x<0) = nan, log(inf) = inf, log(nan) = nan).
DDD_HD double log(double x) {
  const bool special = !(x > 0.0) || !(x < 1.0 / 0.0);  // <= 0, nan or inf
  if (special) {
    if (x == 0.0) return -1.0 / 0.0;
    if (x != x || x > 0.0) return x + x;  // nan or +inf
    return (x - x) / 0.0;                 // negative -> nan
  }
  int k = 0;
  if ((bits_of(x) >> 52) == 0) {  // subnormal: scale into the normal range
    x *= 18014398509481984.0;     // 2^54
    k = -54;
  }
  return log_core(x, k, IeeeDiv{});
}

// atan(x) for all x, its one division through div (numerator in [-1, 1),
// denominator in [1, 4.7) or |x| itself from 2.4375 on).
template <class Div> DDD_HD double atan_core(double x, Div div) {
  const double hi0 = 4.63647609000806093515e-01, lo0 = 2.26987774529616870924e-17;
  const double hi1 = 7.85398163397448278999e-01, lo1 = 3.06161699786838301793e-17;
  const double hi2 = 9.82793723247329054082e-01, lo2 = 1.39033110312309984516e-17;
  const double hi3 = 1.57079632679489655800e+00, lo3 = 6.12323399573676603587e-17;
  const double aT0 = 3.33333333333329318027e-01, aT1 = -1.99999999998764832476e-01,
               aT2 = 1.42857142725034663711e-01, aT3 = -1.11111104054623557880e-01,
               aT4 = 9.09088713343650656196e-02, aT5 = -7.69187620504482999495e-02,
               aT6 = 6.66107313738753120669e-02, aT7 = -5.83357013379057348645e-02,
               aT8 = 4.97687799461593236017e-02, aT9 = -3.65315727442169155270e-02,
               aT10 = 1.62858201153657823623e-02;
  if (x != x) return x + x;
  const bool neg = x < 0.0;
  const double ax = neg ? -x : x;
  // Reduction interval: -1 none, 0 about 1/2, 1 about 1, 2 about 3/2, 3 about inf.
  const int id = ax < 0.4375 ? -1 : ax < 0.6875 ? 0 : ax < 1.1875 ? 1 : ax < 2.4375 ? 2 : 3;
  double num, den;
  if (id == -1) {
    num = ax; den = 1.0;
  } else if (id == 0) {
    num = 2.0 * ax - 1.0; den = 2.0 + ax;
  } else if (id == 1) {
    num = ax - 1.0; den = ax + 1.0;
  } else if (id == 2) {
    num = ax - 1.5; den = 1.0 + 1.5 * ax;
  } else {
    num = -1.0; den = ax;
  }
  const double r = div(num, den);
  const double z = r * r;
  const double w = z * z;
  const double s1 = z * (aT0 + w * (aT2 + w * (aT4 + w * (aT6 + w * (aT8 + w * aT10)))));
  const double s2 = w * (aT1 + w * (aT3 + w * (aT5 + w * (aT7 + w * aT9))));
  const double hi = id == 0 ? hi0 : id == 1 ? hi1 : id == 2 ? hi2 : hi3;
  const double lo = id == 0 ? lo0 : id == 1 ? lo1 : id == 2 ? lo2 : lo3;
  double res = (id < 0) ? r - r * (s1 + s2) : hi - ((r * (s1 + s2) - lo) - r);
  if (ax < 3.7252902984e-09) res = ax;     // |x| < 2^-28: atan(x) = x to 1 ulp
  if (ax >= 7.378697629483821e+19) res = hi3 + lo3;  // |x| >= 2^66
  return neg ? -res : res;
}

DDD_HD double atan(double x) { return atan_core(x, IeeeDiv{}); }

}  // namespace ddd_libm
