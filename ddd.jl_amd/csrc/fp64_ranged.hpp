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
  return __builtin_amdgcn_class(x, 0x100) ? r : __builtin_nan("");  // 0x100: positive normal
}

// ddd_libm::atan (its division's operands are safe for every x; NaN stays NaN).
__device__ __forceinline__ double atan_ranged(double x) {
  return ddd_libm::atan_core(x, RangedDiv{});
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
