// FP64 division, square root, log and atan for the all-pairs tile kernel,
// without the range scaling and special-value fix-ups of the general
// sequences.  The general IEEE sequences on gfx950 are: division,
// div_scale (x2), rcp, two Newton steps (4 fma), mul, residual fma,
// div_fmas, div_fixup; square root, a scale test and ldexp, rsq, two muls,
// seven fmas, ldexp back and a class test for 0 / inf.  When the operands
// are normal, finite and far from the exponent limits, div_scale and
// div_fmas scale nothing and div_fixup returns the quotient unchanged, and
// the square root's scaling is off, so the sequences below -- the same
// operations without those steps -- give the same correctly rounded bits.
// The tile kernel uses them only on tiles whose segment coordinates and
// Burgers vectors lie in the safe range (ranged_ok), and where a zero,
// infinite or NaN operand still occurs these sequences return NaN, never a
// finite wrong value: a tile with a non-finite partial sum is evaluated again
// with the general sequences.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ddd_libm.hpp"

namespace ddd {

// a / b, correctly rounded for normal finite a, b whose quotient and
// exponent gap stay far from the limits (b = 0 or inf gives NaN).
__device__ __forceinline__ double div_ranged(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  double e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  const double q = a * r;
  const double rem = __builtin_fma(-b, q, a);
  return __builtin_fma(rem, r, q);
}

// sqrt(x), correctly rounded for normal finite x >= 2^-767 (0 or inf gives
// NaN, negative x NaN).
__device__ __forceinline__ double sqrt_ranged(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y;
  double h = y * 0.5;
  const double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  return __builtin_fma(d, h, g);
}

struct RangedDiv {
  __device__ __forceinline__ double operator()(double a, double b) const {
    return div_ranged(a, b);
  }
};

// ddd_libm::log for positive normal x (its division's operands are in
// [-0.3, 0.42] and [1.7, 2.42]); NaN for every other x.
__device__ __forceinline__ double log_ranged(double x) {
  const double r = ddd_libm::log_core(x, 0, RangedDiv{});
  // 0x100: positive normal; otherwise the high word of a quiet NaN (any low
  // word makes it NaN), one select instead of two
  const int hi = __builtin_amdgcn_class(x, 0x100) ? __double2hiint(r) : 0x7ff80000;
  return __hiloint2double(hi, __double2loint(r));
}

// hi and lo parts of atan at the reduction points 1/2, 1, 3/2 and infinity
// (ddd_libm::atan_core's constants) at index id + 1, index 0 (no
// reduction) holding zeros: atan_ranged's table, in LDS.  A kernel that
// calls atan_ranged fills it with atan_table_fill before its first call.
namespace {
__shared__ double2 atanHiLo[5];
}

// Lanes 0-4 of the calling wavefront write the table; the caller orders the
// writes before the first atan_ranged (a barrier).
__device__ __forceinline__ void atan_table_fill(int lane) {
  const double hi[5] = {0.0, 4.63647609000806093515e-01, 7.85398163397448278999e-01,
                        9.82793723247329054082e-01, 1.57079632679489655800e+00};
  const double lo[5] = {0.0, 2.26987774529616870924e-17, 3.06161699786838301793e-17,
                        1.39033110312309984516e-17, 6.12323399573676603587e-17};
  if (lane < 5) atanHiLo[lane] = make_double2(hi[lane], lo[lane]);
}

// ddd_libm::atan without branches, the same bits for every x (NaN stays
// NaN).  The reduction interval e = id + 1 (0: none, 1-4: about 1/2, 1,
// 3/2, infinity) comes from integer compares of the high word (the bounds
// 7/16, 11/16, 19/16 and 39/16 have zero low words).  Every interval's
// numerator and denominator is num = S |x| - K, den = K |x| + S with
// (S, K) = (1, 0), (2, 1), (1, 1), (1, 1.5), (0, 1): the same single
// rounding as atan_core's expressions (S |x| and 0 |x| are exact; 1.5 |x|
// is rounded as there).  With hi = lo = 0, hi - ((r s - lo) - r) is
// -(r s - r) = r - r s, atan_core's result without reduction, to the bit
// (rounding is symmetric in sign and x - x = +0 either way).  hi and lo are
// read from atanHiLo, so the interval needs no selects of 64-bit constants.
__device__ __forceinline__ double atan_ranged(double x) {
  const uint32_t hw = (uint32_t)__double2hiint(x);
  const uint32_t hx = hw & 0x7fffffffu;
  const double ax = fabs(x);
  const bool b0 = hx >= 0x3FDC0000u, b1 = hx >= 0x3FE60000u, b2 = hx >= 0x3FF30000u,
             b3 = hx >= 0x40038000u;
  const int e = (int)b0 + (int)b1 + (int)b2 + (int)b3;
  const double2 hl = atanHiLo[e];
  // keep the table read here, ahead of the division and polynomial that
  // hide its latency (nothing moves across this point)
  __builtin_amdgcn_sched_barrier(0);
  const uint32_t sHi = b3 ? 0u : (b0 && !b1) ? 0x40000000u : 0x3FF00000u;
  const uint32_t kHi = !b0 ? 0u : (b2 && !b3) ? 0x3FF80000u : 0x3FF00000u;
  const double S = __hiloint2double((int)sHi, 0), K = __hiloint2double((int)kHi, 0);
  const double num = S * ax - K;
  const double den = K * ax + S;
  const double r = div_ranged(num, den);
  const double z = r * r;
  const double w = z * z;
  const double s1 = z * (3.33333333333329318027e-01 +
                         w * (1.42857142725034663711e-01 +
                              w * (9.09088713343650656196e-02 +
                                   w * (6.66107313738753120669e-02 +
                                        w * (4.97687799461593236017e-02 +
                                             w * 1.62858201153657823623e-02)))));
  const double s2 = w * (-1.99999999998764832476e-01 +
                         w * (-1.11111104054623557880e-01 +
                              w * (-7.69187620504482999495e-02 +
                                   w * (-5.83357013379057348645e-02 +
                                        w * -3.65315727442169155270e-02))));
  double res = hl.x - ((r * (s1 + s2) - hl.y) - r);
  res = ax < 3.7252902984e-09 ? ax : res;
  res = ax >= 7.378697629483821e+19 ? 1.57079632679489655800e+00 + 6.12323399573676603587e-17 : res;
  return __hiloint2double(__double2hiint(res) ^ (int)(hw & 0x80000000u), __double2loint(res));
}

// True when a segment's end points and Burgers vector lie in the range on
// which the tile kernel's ranged sequences are exact: every component zero
// or of magnitude in [2^-300, 2^60] (so that coordinate differences are zero
// or at least 2^-352, and no quantity of the pair formula nears the exponent
// limits).
__device__ __forceinline__ bool ranged_value(double v) {
  const double a = fabs(v);
  return a == 0.0 || (a >= 0x1p-300 && a <= 0x1p60);
}

}  // namespace ddd
