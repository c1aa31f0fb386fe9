// Device math of the force path: the non-singular segment-segment force
// (generic and near-parallel branches), the self force and the field-point
// stress of one segment.  Templated on the scalar so the FP64 path and the
// FP32 far-pair path share one statement of the formulas.
//
// Formulas follow DDD.jl's src/Processing/DislocationForces.jl (pair
// :879-1561, self :414-455) and src/Processing/DislocationFEM.jl:129-299
// (stress), i.e. Cai et al. 2006 / Arsenlis et al. 2007 App. A.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ddd_libm.hpp"
#include "fp64_ranged.hpp"

namespace ddd {

// log/atan of the pair math.  FP64 uses the fixed-sequence functions of
// ddd_libm.hpp (identical bits on host and device when built without FMA
// contraction), FP32 the device library.
__device__ __forceinline__ double ulog(double x) { return ddd_libm::log(x); }
__device__ __forceinline__ double uatan(double x) { return ddd_libm::atan(x); }
__device__ __forceinline__ float ulog(float x) { return logf(x); }
__device__ __forceinline__ float uatan(float x) { return atanf(x); }

template <class T> struct v3 {
  T x, y, z;
};
template <class T> __device__ __forceinline__ v3<T> V(T a, T b, T c) { return v3<T>{a, b, c}; }
template <class T> __device__ __forceinline__ v3<T> operator+(v3<T> a, v3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> __device__ __forceinline__ v3<T> operator-(v3<T> a, v3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> __device__ __forceinline__ v3<T> operator-(v3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> __device__ __forceinline__ v3<T> operator*(T s, v3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> __device__ __forceinline__ v3<T> operator*(v3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> __device__ __forceinline__ T dot(v3<T> a, v3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> __device__ __forceinline__ v3<T> cross(v3<T> a, v3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Pair-kernel constants (segSegForceUnroll, DislocationForces.jl:670-718).
template <class T> struct SegK {
  T aSq, mu4pi, mu8pi, mu8piaSq, mu4pinu, mu4pinuaSq;
};

// One segment as the pair kernel sees it: end nodes, Burgers vector, unit
// tangent and inverse length (the tangent and 1/L are what the reference
// recomputes at the top of every pair call, :880-886).
template <class T> struct Seg {
  v3<T> n1, n2, b, t;
  T tN;
};

// sqrt(eps(Float64)), the branch threshold of :892.
#define DDD_SQRT_EPS 1.4901161193847656e-08

// Unit tangent and 1/L without FMA contraction, so that the branch decision
// of :892 sees the same bits as an IEEE restatement of the reference.
__device__ __forceinline__ void unit_tangent(v3<double> n1, v3<double> n2, v3<double>& t, double& tN) {
  double dx = __dsub_rn(n2.x, n1.x), dy = __dsub_rn(n2.y, n1.y), dz = __dsub_rn(n2.z, n1.z);
  double ss = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
  tN = __ddiv_rn(1.0, __dsqrt_rn(ss));
  t = V(__dmul_rn(dx, tN), __dmul_rn(dy, tN), __dmul_rn(dz, tN));
}

// 1 - (t1.t2)^2, uncontracted (see unit_tangent).
__device__ __forceinline__ double one_minus_c2(v3<double> t1, v3<double> t2, double& c) {
  c = __dadd_rn(__dadd_rn(__dmul_rn(t1.x, t2.x), __dmul_rn(t1.y, t2.y)), __dmul_rn(t1.z, t2.z));
  return __dsub_rn(1.0, __dmul_rn(c, c));
}

// Approximate-path helpers (FAST below): a reciprocal and a quotient from
// the hardware reciprocal refined by two Newton steps, and ddd_libm's log and
// atan (the same fdlibm reductions and polynomials) with those divisions and
// without the special-case branches, for positive finite (log) and finite
// (atan) arguments.  They differ from the IEEE quotient and from ddd_libm by
// rounding-level amounts; the branch-free code keeps each FP64 pair one
// long basic block.
__device__ __forceinline__ double frcp(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  return fma(r, fma(-d, r, 1.0), r);
}
__device__ __forceinline__ double fdiv(double n, double d) {
  const double r = frcp(d);
  const double q = n * r;
  return fma(fma(-d, q, n), r, q);
}
__device__ __forceinline__ double flog(double x) {
  // Every candidate is computed and then selected, so the code has no branch.
  const bool sub = (__double_as_longlong(x) >> 52) == 0;  // subnormal: scale by 2^54
  const double xs = x * 18014398509481984.0;
  x = sub ? xs : x;
  const uint64_t u = (uint64_t)__double_as_longlong(x);
  uint32_t hx = (uint32_t)(u >> 32) + (0x3ff00000u - 0x3fe6a09eu);
  const int k = (int)(hx >> 20) - 0x3ff - (sub ? 54 : 0);
  hx = (hx & 0x000fffffu) + 0x3fe6a09eu;
  const double f = __longlong_as_double((long long)(((uint64_t)hx << 32) | (u & 0xffffffffull))) - 1.0;
  const double hfsq = 0.5 * f * f;
  const double s = fdiv(f, 2.0 + f);
  const double z = s * s, w = z * z;
  const double t1 = w * (3.999999999940941908e-01 +
                         w * (2.222219843214978396e-01 + w * 1.531383769920937332e-01));
  const double t2 = z * (6.666666666666735130e-01 +
                         w * (2.857142874366239149e-01 +
                              w * (1.818357216161805012e-01 + w * 1.479819860511658591e-01)));
  const double dk = (double)k;
  return s * (hfsq + (t2 + t1)) + dk * 1.90821492927058770002e-10 - hfsq + f +
         dk * 6.93147180369123816490e-01;
}
__device__ __forceinline__ double fatan(double x) {
  const double ax = fabs(x);
  // reduction intervals about 1/2, 1, 3/2 and infinity (none below 7/16)
  const bool b0 = ax >= 0.4375, b1 = ax >= 0.6875, b2 = ax >= 1.1875, b3 = ax >= 2.4375;
  const double n0 = 2.0 * ax - 1.0, d0 = 2.0 + ax;
  const double n1 = ax - 1.0, d1 = ax + 1.0;
  const double n2 = ax - 1.5, d2 = 1.0 + 1.5 * ax;
  double num = ax, den = 1.0, hi = 0.0, lo = 0.0;
  num = b0 ? n0 : num;  den = b0 ? d0 : den;
  hi = b0 ? 4.63647609000806093515e-01 : hi;  lo = b0 ? 2.26987774529616870924e-17 : lo;
  num = b1 ? n1 : num;  den = b1 ? d1 : den;
  hi = b1 ? 7.85398163397448278999e-01 : hi;  lo = b1 ? 3.06161699786838301793e-17 : lo;
  num = b2 ? n2 : num;  den = b2 ? d2 : den;
  hi = b2 ? 9.82793723247329054082e-01 : hi;  lo = b2 ? 1.39033110312309984516e-17 : lo;
  num = b3 ? -1.0 : num;  den = b3 ? ax : den;
  hi = b3 ? 1.57079632679489655800e+00 : hi;  lo = b3 ? 6.12323399573676603587e-17 : lo;
  const double r = fdiv(num, den);
  const double z = r * r, w = z * z;
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
  const double rs = r * (s1 + s2);
  const double reduced = hi - ((rs - lo) - r), direct = r - rs;
  double res = b0 ? reduced : direct;
  res = ax < 3.7252902984e-09 ? ax : res;
  res = ax >= 7.378697629483821e+19 ? 1.57079632679489655800e+00 + 6.12323399573676603587e-17 : res;
  return x < 0.0 ? -res : res;
}

// Sum over the four integration corners of the 19 antiderivatives of
// SegSegInteg (:1480-1561).  Terms shared by all corners (a^2 + d^2 (1-c^2)
// and its 1/sqrt) are evaluated once.  Corners are combined in the
// reference's order: (x1,y1) - (x1,y2) - (x2,y1) + (x2,y2).
template <class T> struct Integ19 {
  T v[19];
};

// FAST (approximate paths only, T = double): the three reciprocals 1/Ra,
// 1/(Ra t1), 1/(Ra t2) from one reciprocal, the division by a^2 + d^2 (1-c^2)
// as a product with its reciprocal aSq_dSqI, flog and fatan; other rounding
// than the reference's.  RANGED (T = double, not FAST): the same operations
// as the default, their divisions, square roots, log and atan through the
// ranged sequences of fp64_ranged.hpp (the same bits on the tiles that use
// them).
template <class T, bool FAST = false, bool RANGED = false>
__device__ __forceinline__ void corner_acc(T aSq_dSq, T aSq_dSqI, T den, T c, T cSq, T omcSqI, T x,
                                           T y, T sgn, bool first, Integ19<T>& I) {
  T xSq = x * x;
  T ySq = y * y;
  const T RaSq = aSq_dSq + xSq + ySq + T(2) * x * y * c;
  T Ra;
  if constexpr (RANGED) Ra = sqrt_ranged(RaSq);
  else Ra = sqrt(RaSq);
  T Ra_Rd_t1 = Ra + y + x * c;
  T Ra_Rd_t2 = Ra + x + y * c;
  T RaInv, R1I, R2I, log1, log2;
  if constexpr (FAST) {
    const T t12 = Ra_Rd_t1 * Ra_Rd_t2;
    const T inv = frcp(Ra * t12);
    RaInv = t12 * inv;
    R1I = Ra_Rd_t2 * inv;
    R2I = Ra_Rd_t1 * inv;
    log1 = flog(Ra_Rd_t1);
    log2 = flog(Ra_Rd_t2);
  } else if constexpr (RANGED) {
    RaInv = div_ranged(T(1), Ra);
    R1I = div_ranged(RaInv, Ra_Rd_t1);
    R2I = div_ranged(RaInv, Ra_Rd_t2);
    log1 = log_ranged(Ra_Rd_t1);
    log2 = log_ranged(Ra_Rd_t2);
  } else {
    RaInv = T(1) / Ra;
    R1I = RaInv / Ra_Rd_t1;
    R2I = RaInv / Ra_Rd_t2;
    log1 = ulog(Ra_Rd_t1);
    log2 = ulog(Ra_Rd_t2);
  }
  T xlog1 = x * log1;
  T ylog2 = y * log2;
  T xR1I = x * R1I;
  T xSqR1I = x * xR1I;
  T yR2I = y * R2I;
  T ySqR2I = y * yR2I;

  const T atnArg = (T(1) + c) * (Ra + x + y) * den;
  T i1;
  if constexpr (FAST) i1 = T(-2) * den * fatan(atnArg);
  else if constexpr (RANGED) i1 = T(-2) * den * atan_ranged(atnArg);
  else i1 = T(-2) * den * uatan(atnArg);
  T c_1 = aSq_dSq * i1;
  T c_5_6 = (c * Ra - c_1) * omcSqI;
  T i2 = (c * log2 - log1) * omcSqI;
  T i3 = (c * log1 - log2) * omcSqI;
  T i4 = (c * c_1 - Ra) * omcSqI;
  T i5 = ylog2 + c_5_6;
  T i6 = xlog1 + c_5_6;
  T c_11_12 = i1 - c * RaInv;
  T c_15_18 = c * xR1I - RaInv;
  T x_13_14 = x * c_15_18;
  T c_19 = c * yR2I - RaInv;
  T y_13_14 = y * c_19;
  T c_16 = log2 - (x - c * y) * RaInv - cSq * ySqR2I;
  T z_15_18 = y * c_16;
  T c_17_19 = log1 - (y - c * x) * RaInv - cSq * xSqR1I;
  T c15_18_19 = T(2) * i4;

  T r[19];
  r[0] = i1;
  r[1] = i2;
  r[2] = i3;
  r[3] = i4;
  r[4] = i5;
  r[5] = i6;
  if constexpr (FAST) r[6] = (i1 - xR1I - yR2I) * aSq_dSqI;
  else if constexpr (RANGED) r[6] = div_ranged(i1 - xR1I - yR2I, aSq_dSq);
  else r[6] = (i1 - xR1I - yR2I) / aSq_dSq;
  r[7] = (R1I - c * R2I) * omcSqI;
  r[8] = (R2I - c * R1I) * omcSqI;
  r[9] = (RaInv - c * (xR1I + yR2I + i1)) * omcSqI;
  r[10] = (xR1I + cSq * yR2I + c_11_12) * omcSqI;
  r[11] = (yR2I + cSq * xR1I + c_11_12) * omcSqI;
  r[12] = (i3 - x_13_14 + c * (y_13_14 - i2)) * omcSqI;
  r[13] = (i2 - y_13_14 + c * (x_13_14 - i3)) * omcSqI;
  r[14] = (i5 - z_15_18 + c * (xSq * c_15_18 - c15_18_19)) * omcSqI;
  r[15] = (xSqR1I + c * c_16 + T(2) * i2) * omcSqI;
  r[16] = (ySqR2I + c * c_17_19 + T(2) * i3) * omcSqI;
  r[17] = (c15_18_19 - xSq * c_15_18 + c * (z_15_18 - i5)) * omcSqI;
  r[18] = (c15_18_19 - ySq * c_19 + c * (x * c_17_19 - i6)) * omcSqI;
#pragma unroll
  for (int k = 0; k < 19; ++k) I.v[k] = first ? r[k] : I.v[k] + sgn * r[k];
}

// Generic pair (:893-1146).  W12: forces on segment 1's nodes (F1, F2);
// W34: forces on segment 2's nodes (F3, F4).  c and omcSq come from the
// caller (the branch test already computed them).  FAST: corner_acc's
// shared reciprocals (approximate paths only).  RANGED: the ranged
// sequences of fp64_ranged.hpp (see corner_acc).
template <class T, bool W12, bool W34, bool FAST = false, bool RANGED = false>
__device__ __forceinline__ void pair_generic(const SegK<T>& K, const Seg<T>& s1, const Seg<T>& s2,
                                             T c, T omcSq, v3<T>& F1, v3<T>& F2, v3<T>& F3,
                                             v3<T>& F4) {
  const v3<T> t1 = s1.t, t2 = s2.t, b1 = s1.b, b2 = s2.b;
  const T cSq = c * c;
  T omcSqI;
  if constexpr (FAST) omcSqI = frcp(omcSq);
  else if constexpr (RANGED) omcSqI = div_ranged(T(1), omcSq);
  else omcSqI = T(1) / omcSq;

  const v3<T> t2ct1 = cross(t2, t1);
  const v3<T> t1ct2 = -t2ct1;
  const v3<T> b2ct2 = cross(b2, t2);
  const v3<T> b1ct1 = cross(b1, t1);
  const T t2db2 = dot(t2, b2), t2db1 = dot(t2, b1), t1db2 = dot(t1, b2), t1db1 = dot(t1, b1);
  const T t2ct1db2 = dot(t2ct1, b2), t1ct2db1 = dot(t1ct2, b1);
  const T b1ct1db2 = dot(b1ct1, b2), b2ct2db1 = dot(b2ct2, b1);
  const v3<T> t2ct1ct2 = t1 - c * t2;
  const v3<T> t1ct2ct1 = t2 - c * t1;
  const T t2ct1cb1dt1 = t2db1 - t1db1 * c;
  const T t1ct2cb2dt2 = t1db2 - t2db2 * c;
  const T t2ct1cb1db2 = t2db1 * t1db2 - t1db1 * t2db2;

  const v3<T> R1 = s2.n1 - s1.n1;
  const v3<T> R2 = s2.n2 - s1.n2;
  const T d = dot(R2, t2ct1) * omcSqI;
  const T mu4pid = K.mu4pi * d;
  const T mu4pinud = K.mu4pinu * d;
  const T mu4pinudSq = mu4pinud * d;
  const T mu4pinudCu = mu4pinudSq * d;
  const T mu4pinuaSqd = K.mu4pinuaSq * d;
  const T mu8piaSqd = K.mu8piaSq * d;
  const T lim11 = dot(R1, t1), lim12 = dot(R1, t2), lim21 = dot(R2, t1), lim22 = dot(R2, t2);
  const T x1 = (lim12 - c * lim11) * omcSqI;
  const T x2 = (lim22 - c * lim21) * omcSqI;
  const T y1 = (lim11 - c * lim12) * omcSqI;
  const T y2 = (lim21 - c * lim22) * omcSqI;

  const T aSq_dSq = K.aSq + d * d * omcSq;
  T den;
  if constexpr (FAST) den = frcp(sqrt(omcSq * aSq_dSq));
  else if constexpr (RANGED) den = div_ranged(T(1), sqrt_ranged(omcSq * aSq_dSq));
  else den = T(1) / sqrt(omcSq * aSq_dSq);
  T aSq_dSqI = T(0);
  if constexpr (FAST) aSq_dSqI = frcp(aSq_dSq);
  Integ19<T> I;
  corner_acc<T, FAST, RANGED>(aSq_dSq, aSq_dSqI, den, c, cSq, omcSqI, x1, y1, T(1), true, I);
  corner_acc<T, FAST, RANGED>(aSq_dSq, aSq_dSqI, den, c, cSq, omcSqI, x1, y2, T(-1), false, I);
  corner_acc<T, FAST, RANGED>(aSq_dSq, aSq_dSqI, den, c, cSq, omcSqI, x2, y1, T(-1), false, I);
  corner_acc<T, FAST, RANGED>(aSq_dSq, aSq_dSqI, den, c, cSq, omcSqI, x2, y2, T(1), false, I);
  const T* J = I.v - 1;  // 1-based, as the formulas are written

  if (W12) {  // segment 1, nodes 1-2: :953-1048
    const v3<T> t1cb2ct1 = b2 - t1db2 * t1;
    const v3<T> b2ct2ct1 = t1db2 * t2 - c * b2;
    const T tq = t1db2 * t2db1 + t2ct1cb1db2;
    const v3<T> P1 = tq * t2ct1;
    const v3<T> P2 = b1ct1 * t1ct2cb2dt2;
    const v3<T> P3 = t1ct2 * b1ct1db2 - t1cb2ct1 * t2db1;
    const v3<T> P4 = -b1ct1 * t2ct1db2;
    const v3<T> P5 = b2ct2ct1 * t2db1 - t1ct2 * b2ct2db1;
    v3<T> W[11];
    W[0] = mu4pid * P1 - mu4pinud * P2 + (mu4pinud * t1ct2db1) * b2ct2ct1 +
           (mu4pinud * b2ct2db1) * t1ct2ct1;
    W[1] = K.mu4pi * P5 - (K.mu4pinu * t2db1) * b2ct2ct1 + (K.mu4pinu * b2ct2db1) * t1ct2;
    W[2] = -(K.mu4pinu * t1db1) * b2ct2ct1 + K.mu4pi * P3 - K.mu4pinu * P4;
    W[3] = mu8piaSqd * P1 - mu4pinuaSqd * P2 - (mu4pinudCu * t1ct2cb2dt2 * t1ct2db1) * t1ct2ct1;
    W[4] = K.mu8piaSq * P5 + (mu4pinudSq * t1ct2cb2dt2 * t2db1) * t1ct2ct1 -
           (mu4pinudSq * t1ct2cb2dt2 * t1ct2db1) * t1ct2;
    W[5] = K.mu8piaSq * P3 - K.mu4pinuaSq * P4 +
           (mu4pinudSq * (t2ct1db2 * t1ct2db1 + t1ct2cb2dt2 * t1db1)) * t1ct2ct1;
    W[6] = (mu4pinud * (t1ct2cb2dt2 * t1db1 + t2ct1db2 * t1ct2db1)) * t1ct2 -
           (mu4pinud * t2ct1db2 * t2db1) * t1ct2ct1;
    W[7] = (mu4pinud * t1ct2cb2dt2 * t2db1) * t1ct2;
    W[8] = -(mu4pinud * t2ct1db2 * t1db1) * t1ct2ct1;
    const v3<T> tv = (K.mu4pinu * t2ct1db2) * t1ct2;
    W[9] = -tv * t2db1;
    W[10] = -tv * t1db1;
    // Fint pairs (a, b): node 1 uses J[a] - y2 J[b], node 2 uses y1 J[b] - J[a].
    const int A[11] = {3, 4, 6, 9, 10, 12, 14, 13, 17, 15, 19};
    const int B[11] = {1, 2, 3, 7, 8, 9, 10, 11, 12, 13, 14};
    v3<T> a1 = W[0] * (J[A[0]] - y2 * J[B[0]]);
    v3<T> a2 = W[0] * (y1 * J[B[0]] - J[A[0]]);
#pragma unroll
    for (int k = 1; k < 11; ++k) {
      a1 = a1 + W[k] * (J[A[k]] - y2 * J[B[k]]);
      a2 = a2 + W[k] * (y1 * J[B[k]] - J[A[k]]);
    }
    F1 = a1 * s1.tN;
    F2 = a2 * s1.tN;
  }
  if (W34) {  // segment 2, nodes 3-4: :1050-1145
    const v3<T> t2cb1ct2 = b1 - t2db1 * t2;
    const v3<T> b1ct1ct2 = t2db1 * t1 - c * b1;
    const T tq = t2db1 * t1db2 + t2ct1cb1db2;
    const v3<T> P1 = tq * t1ct2;
    const v3<T> P2 = b2ct2 * t2ct1cb1dt1;
    const v3<T> P3 = t2ct1 * b1ct1db2 - b1ct1ct2 * t1db2;
    const v3<T> P5 = t2cb1ct2 * t1db2 - t2ct1 * b2ct2db1;
    const v3<T> P6 = b2ct2 * t1ct2db1;
    v3<T> W[11];
    W[0] = mu4pid * P1 - mu4pinud * P2 + (mu4pinud * t2ct1db2) * b1ct1ct2 +
           (mu4pinud * b1ct1db2) * t2ct1ct2;
    W[1] = (K.mu4pinu * t2db2) * b1ct1ct2 + K.mu4pi * P5 - K.mu4pinu * P6;
    W[2] = K.mu4pi * P3 + (K.mu4pinu * t1db2) * b1ct1ct2 - (K.mu4pinu * b1ct1db2) * t2ct1;
    W[3] = mu8piaSqd * P1 - mu4pinuaSqd * P2 - (mu4pinudCu * t2ct1cb1dt1 * t2ct1db2) * t2ct1ct2;
    W[4] = K.mu8piaSq * P5 - K.mu4pinuaSq * P6 -
           (mu4pinudSq * (t1ct2db1 * t2ct1db2 + t2ct1cb1dt1 * t2db2)) * t2ct1ct2;
    W[5] = K.mu8piaSq * P3 - (mu4pinudSq * t2ct1cb1dt1 * t1db2) * t2ct1ct2 +
           (mu4pinudSq * t2ct1cb1dt1 * t2ct1db2) * t2ct1;
    W[6] = (mu4pinud * (t2ct1cb1dt1 * t2db2 + t1ct2db1 * t2ct1db2)) * t2ct1 -
           (mu4pinud * t1ct2db1 * t1db2) * t2ct1ct2;
    W[7] = -(mu4pinud * t1ct2db1 * t2db2) * t2ct1ct2;
    W[8] = (mu4pinud * t2ct1cb1dt1 * t1db2) * t2ct1;
    const v3<T> tv = (K.mu4pinu * t1ct2db1) * t2ct1;
    W[9] = tv * t2db2;
    W[10] = tv * t1db2;
    // node 3 uses x2 J[b] - J[a], node 4 uses J[a] - x1 J[b].
    const int A[11] = {2, 5, 4, 8, 11, 10, 13, 16, 14, 18, 15};
    const int B[11] = {1, 2, 3, 7, 8, 9, 10, 11, 12, 13, 14};
    v3<T> a3 = W[0] * (x2 * J[B[0]] - J[A[0]]);
    v3<T> a4 = W[0] * (J[A[0]] - x1 * J[B[0]]);
#pragma unroll
    for (int k = 1; k < 11; ++k) {
      a3 = a3 + W[k] * (x2 * J[B[k]] - J[A[k]]);
      a4 = a4 + W[k] * (J[A[k]] - x1 * J[B[k]]);
    }
    F3 = a3 * s2.tN;
    F4 = a4 * s2.tN;
  }
}

// Near-parallel antiderivatives (:1444-1478), summed over the 4 corners.
template <class T>
__device__ __forceinline__ void par_corner_acc(T aSq_dSq, T aSq_dSqI, T x, T y, T sgn, bool first,
                                               T* I) {
  T xpy = x + y;
  T xmy = x - y;
  T Ra = sqrt(aSq_dSq + xpy * xpy);
  T RaInv = T(1) / Ra;
  T L = ulog(Ra + xpy);
  T tmp = xmy * Ra * aSq_dSqI;
  T i10 = -aSq_dSqI * xpy * RaInv;
  T r[12] = {Ra * aSq_dSqI,
             T(-0.5) * (L - tmp),
             T(-0.5) * (L + tmp),
             -L,
             y * L - Ra,
             x * L - Ra,
             aSq_dSqI * (T(2) * aSq_dSqI * Ra - RaInv),
             aSq_dSqI * (tmp - x * RaInv),
             -aSq_dSqI * (tmp + y * RaInv),
             i10,
             RaInv - y * i10,
             RaInv - x * i10};
#pragma unroll
  for (int k = 0; k < 12; ++k) I[k] = first ? r[k] : I[k] + sgn * r[k];
}

// One side of calcParSegSegForce (:1204-1318 for side 1, :1320-1434 for side
// 2): forces on the nodes of segment A (a1 -> a2, Burgers bA, unit tangent tA,
// 1/L tNA) from segment B (b1 -> b2, Burgers bB) projected on A's line and
// tilted by the critical angle.  Returns the forces on A's first and second
// node in (Fa1, Fa2), and the two tilted pieces of B for the corrective
// generic calls in (c1a, c1b), (c2a, c2b) with doCorr set when they apply.
// The first steps of the near-parallel branch (:1203-1213): the other
// segment p1 -> p2 projected onto A's line (unit tangent tA) and tilted by
// the critical angle to p1m -> p2m; magDiffSq0 = |rejection|^2, magDiff its
// root.  par_side and the corrective pass share it, so both see the same
// bits.
template <class T>
__device__ __forceinline__ void par_tilt(v3<T> tA, v3<T> p1, v3<T> p2, v3<T>& p1m, v3<T>& p2m,
                                         T& magDiffSq0, T& magDiff) {
  const T se = T(DDD_SQRT_EPS);
  const T hCot = sqrt((T(1) - se * T(1.01)) / (se * T(1.01))) / T(2);
  const T tmp = dot(p2 - p1, tA);
  const v3<T> q = p1 + tmp * tA;
  const v3<T> diff = p2 - q;
  magDiffSq0 = dot(diff, diff);
  magDiff = sqrt(magDiffSq0);
  const v3<T> tv1 = T(0.5) * diff;
  const v3<T> tv2 = (hCot * magDiff) * tA;
  p1m = p1 + tv1 + tv2;
  p2m = q + tv1 - tv2;
}

template <class T>
__device__ __forceinline__ void par_side(const SegK<T>& K, v3<T> a1, v3<T> a2, v3<T> tA, T tNA,
                                         v3<T> bA, v3<T> p1, v3<T> p2, v3<T> bB, bool sideOne,
                                         v3<T>& Fa1, v3<T>& Fa2, v3<T>& c1a, v3<T>& c1b,
                                         v3<T>& c2a, v3<T>& c2b, bool& doCorr) {
  const T se = T(DDD_SQRT_EPS);
  // Project the other segment (p1 -> p2) on A's line and tilt it.
  v3<T> p1m, p2m;
  T magDiffSq0, magDiff;
  par_tilt(tA, p1, p2, p1m, p2m, magDiffSq0, magDiff);

  // Side 1 measures from A's first node to the tilted other segment; side 2
  // from the tilted segment (here the roles of "this" and "other" swap in
  // the reference's variable names, but the arithmetic is the same shape).
  const v3<T> R = sideOne ? (p1m - a1) : (a1 - p1m);
  const T Rdt = dot(R, tA);
  const v3<T> nd = R - Rdt * tA;
  const T dSq = dot(nd, nd);
  const T aSq_dSq = K.aSq + dSq;
  const T aSq_dSqI = T(1) / aSq_dSq;
  T x1, x2, y1, y2;
  if (sideOne) {
    x1 = dot(p1m, tA);
    x2 = dot(p2m, tA);
    y1 = -dot(a1, tA);
    y2 = -dot(a2, tA);
  } else {
    x1 = dot(a1, tA);
    x2 = dot(a2, tA);
    y1 = -dot(p1m, tA);
    y2 = -dot(p2m, tA);
  }
  const T tAdbB = dot(tA, bB);
  const T tAdbA = dot(tA, bA);
  const T nddbA = dot(nd, bA);
  const v3<T> bBctA = cross(bB, tA);
  const v3<T> bActA = cross(bA, tA);
  const v3<T> ndctA = cross(nd, tA);
  const T bBctAdbA = dot(bBctA, bA);
  const T bBctAdnd = dot(bBctA, nd);
  const v3<T> bBctActA = tAdbB * tA - bB;

  T I[12];
  par_corner_acc(aSq_dSq, aSq_dSqI, x1, y1, T(1), true, I);
  par_corner_acc(aSq_dSq, aSq_dSqI, x1, y2, T(-1), false, I);
  par_corner_acc(aSq_dSq, aSq_dSqI, x2, y1, T(-1), false, I);
  par_corner_acc(aSq_dSq, aSq_dSqI, x2, y2, T(1), false, I);
  const T* J = I - 1;

  const v3<T> w1 = (tAdbA * tAdbB) * nd;
  const v3<T> w2 = bBctAdnd * bActA;
  const v3<T> U1 = K.mu4pinu * (nddbA * bBctActA + bBctAdbA * ndctA - w2) - K.mu4pi * w1;
  const v3<T> U2 = ((K.mu4pinu - K.mu4pi) * tAdbA) * bBctActA;
  const v3<T> U3 = -K.mu8piaSq * w1 - K.mu4pinuaSq * w2 - (K.mu4pinu * bBctAdnd * nddbA) * ndctA;
  const v3<T> U4 = -(K.mu8piaSq * tAdbA) * bBctActA - (K.mu4pinu * bBctAdnd * tAdbA) * ndctA;
  if (sideOne) {
    // node 2: J[3k+3] - y1 J[3k+1]; node 1: y2 J[3k+1] - J[3k+3]
    Fa2 = (U1 * (J[3] - y1 * J[1]) + U2 * (J[6] - y1 * J[4]) + U3 * (J[9] - y1 * J[7]) +
           U4 * (J[12] - y1 * J[10])) * tNA;
    Fa1 = (U1 * (y2 * J[1] - J[3]) + U2 * (y2 * J[4] - J[6]) + U3 * (y2 * J[7] - J[9]) +
           U4 * (y2 * J[10] - J[12])) * tNA;
  } else {
    // node 4 (second): J[3k+2] - x1 J[3k+1]; node 3 (first): x2 J[3k+1] - J[3k+2]
    Fa2 = (U1 * (J[2] - x1 * J[1]) + U2 * (J[5] - x1 * J[4]) + U3 * (J[8] - x1 * J[7]) +
           U4 * (J[11] - x1 * J[10])) * tNA;
    Fa1 = (U1 * (x2 * J[1] - J[2]) + U2 * (x2 * J[4] - J[5]) + U3 * (x2 * J[7] - J[8]) +
           U4 * (x2 * J[10] - J[11])) * tNA;
  }
  const T magDiffSq = sideOne ? magDiffSq0 : magDiff * magDiff;
  doCorr = magDiffSq > se * (dot(p1m, p1m) + dot(p2m, p2m));
  c1a = p1;
  c1b = p1m;
  c2a = p2m;
  c2b = p2;
}

// Stress constants of the field-point formula (DislocationFEM.jl:96-101).
template <class T> struct SigK {
  T a2, mu4pi, mu4pinu, nuMu4pinu, a2mu8pi;
};

// Non-singular stress at point x of a straight segment from p to p + d
// with unit tangent t and Burgers vector b, accumulated into sg
// (xx, yy, zz, xy, xz, yz).  x is given relative to p (rx).  The terms follow
// DislocationFEM.jl:142-299; t x b and the t-only products are the caller's
// per-segment invariants.
template <class T>
__device__ __forceinline__ void seg_stress(const SigK<T>& K, v3<T> rx, v3<T> d, v3<T> t,
                                           v3<T> b, v3<T> txb, T sg[6]) {
  const T Rdt = dot(rx, t);
  const v3<T> nd = rx - Rdt * t;
  const T d2 = dot(nd, nd);
  const T s1 = -Rdt;
  const T s2 = -dot(rx - d, t);
  const T a2_d2 = K.a2 + d2;
  const T a2d2inv = T(1) / a2_d2;
  T Rainv = T(1) / sqrt(a2_d2 + s1 * s1);
  T Ra3inv = Rainv * Rainv * Rainv;
  T sRa3inv = s1 * Ra3inv;
  const T s03a = s1 * Rainv * a2d2inv;
  const T s13a = -Rainv;
  const T s05a = (T(2) * s03a + sRa3inv) * a2d2inv;
  const T s15a = -Ra3inv;
  const T s25a = s03a - sRa3inv;
  Rainv = T(1) / sqrt(a2_d2 + s2 * s2);
  Ra3inv = Rainv * Rainv * Rainv;
  sRa3inv = s2 * Ra3inv;
  const T s03b = s2 * Rainv * a2d2inv;
  const T s05b = (T(2) * s03b + sRa3inv) * a2d2inv;
  const T s03 = s03b - s03a, s13 = -Rainv - s13a, s05 = s05b - s05a, s15 = -Ra3inv - s15a,
          s25 = (s03b - sRa3inv) - s25a;
  const v3<T> dxb = cross(nd, b);
  const T common = K.mu4pinu * dot(dxb, t);
  const T tt[3] = {t.x, t.y, t.z}, dd[3] = {nd.x, nd.y, nd.z};
  const T xb[3] = {txb.x, txb.y, txb.z}, db[3] = {dxb.x, dxb.y, dxb.z};
  const int PP[6] = {0, 1, 2, 0, 0, 1}, QQ[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const int a = PP[q], c = QQ[q];
    const bool dg = q < 3;
    const T dmd = dd[a] * dd[c];
    const T tmt = tt[a] * tt[c];
    const T tmd = dg ? T(2) * tt[a] * dd[a] : tt[a] * dd[c] + tt[c] * dd[a];
    const T tmtxb = dg ? T(2) * tt[a] * xb[a] : tt[a] * xb[c] + tt[c] * xb[a];
    const T dmtxb = dg ? T(2) * dd[a] * xb[a] : dd[a] * xb[c] + dd[c] * xb[a];
    const T tmdxb = dg ? T(2) * tt[a] * db[a] : tt[a] * db[c] + tt[c] * db[a];
    const T I03 = (dg ? common : T(0)) + K.mu4pinu * dmtxb - K.mu4pi * tmdxb;
    const T I13 = -K.nuMu4pinu * tmtxb;
    const T I05 = common * ((dg ? K.a2 : T(0)) + dmd) - K.a2mu8pi * tmdxb;
    const T I15 = K.a2mu8pi * tmtxb - common * tmd;
    const T I25 = common * tmt;
    sg[q] += I03 * s03 + I13 * s13 + I05 * s05 + I15 * s15 + I25 * s25;
  }
}

// The stress of seg_stress at rx (relative to p) applied to a vector ba,
// sigma(rx) ba, without forming the six components: with the segment's
// length L = d.t, its field regrouped as
//   sigma = delta cD + t (x) u + u (x) t + n (x) v + v (x) n
// (n the rejection of rx from the line, u and v vectors of the point), whose
// product with ba takes four dot products.  The same mathematics as
// seg_stress in about half the operations, rounded otherwise: for the
// approximate paths (far_side).
template <class T>
__device__ __forceinline__ T q_rcp(T x) { return T(1) / x; }
template <class T>
__device__ __forceinline__ T q_rsqrt(T x) { return T(1) / sqrt(x); }
// FP32: the hardware reciprocal and reciprocal square root (about 1 ulp)
// without the library forms' denormal range scaling; their arguments here
// are at least a^2 (refused outside [2^-60, 2^60] on the host) or half a
// segment length, far from the denormal range.
template <>
__device__ __forceinline__ float q_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
template <>
__device__ __forceinline__ float q_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }

template <class T>
__device__ __forceinline__ v3<T> seg_stress_times(const SigK<T>& K, v3<T> rx, T L, v3<T> t, v3<T> b,
                                                  v3<T> txb, T tba, v3<T> ba) {
  const T Rdt = dot(rx, t);
  const v3<T> nd = rx - Rdt * t;
  const T d2 = dot(nd, nd);
  const T s1 = -Rdt;
  const T s2 = L - Rdt;
  const T a2_d2 = K.a2 + d2;
  const T a2d2inv = q_rcp(a2_d2);
  T Rainv = q_rsqrt(a2_d2 + s1 * s1);
  T Ra3inv = Rainv * Rainv * Rainv;
  T sRa3inv = s1 * Ra3inv;
  const T s03a = s1 * Rainv * a2d2inv;
  const T s13a = -Rainv;
  const T s05a = (T(2) * s03a + sRa3inv) * a2d2inv;
  const T s15a = -Ra3inv;
  const T s25a = s03a - sRa3inv;
  Rainv = q_rsqrt(a2_d2 + s2 * s2);
  Ra3inv = Rainv * Rainv * Rainv;
  sRa3inv = s2 * Ra3inv;
  const T s03b = s2 * Rainv * a2d2inv;
  const T s05b = (T(2) * s03b + sRa3inv) * a2d2inv;
  const T s03 = s03b - s03a, s13 = -Rainv - s13a, s05 = s05b - s05a, s15 = -Ra3inv - s15a,
          s25 = (s03b - sRa3inv) - s25a;
  const v3<T> dxb = cross(nd, b);
  const T common = K.mu4pinu * dot(dxb, t);
  // coefficients of delta, n(x)n, n(x)txb, t(x)dxb, t(x)txb, t(x)n, t(x)t in
  // the integrals I03 s03 + I13 s13 + I05 s05 + I15 s15 + I25 s25
  const T cD = common * (s03 + K.a2 * s05);
  const T cnn = common * s05;
  const T cnx = K.mu4pinu * s03;
  const T cty = -(K.mu4pi * s03 + K.a2mu8pi * s05);
  const T ctx = K.a2mu8pi * s15 - K.nuMu4pinu * s13;
  const T ctn = -common * s15;
  const T ctt = common * s25;
  const v3<T> u = cty * dxb + ctx * txb + ctn * nd + (T(0.5) * ctt) * t;
  const v3<T> v = (T(0.5) * cnn) * nd + cnx * txb;
  return cD * ba + dot(u, ba) * t + tba * u + dot(v, ba) * nd + dot(nd, ba) * v;
}

// Far-pair nodal forces by two-point Gauss-Legendre quadrature (no reference
// counterpart; the mixed path's formulation for well-separated pairs): on
// segment A, F_A1 = L_A sum_q w_q (1 - u_q) (sigma_B(x_q) b_A) x t_A and
// F_A2 likewise with u_q, sigma_B the non-singular stress of B.  Positions
// enter as double differences rounded once to T, so T = float only carries
// lengths of the pair's own scale.  Against the FP64 analytic force the
// per-node error is ~3e-7 RMS on 8,192-segment networks (quadrature-bound).
template <class T>
__device__ __forceinline__ void far_side(const SigK<T>& K, const Seg<double>& A,
                                         const Seg<double>& B, v3<double>& FA1,
                                         v3<double>& FA2) {
  const T g0 = T(0.21132486540518711775), g1 = T(0.78867513459481288225);  // (1 -+ 1/sqrt3)/2
  const v3<T> r0 = V(T(A.n1.x - B.n1.x), T(A.n1.y - B.n1.y), T(A.n1.z - B.n1.z));
  const v3<T> dA = V(T(A.n2.x - A.n1.x), T(A.n2.y - A.n1.y), T(A.n2.z - A.n1.z));
  const v3<T> dB = V(T(B.n2.x - B.n1.x), T(B.n2.y - B.n1.y), T(B.n2.z - B.n1.z));
  const v3<T> tB = V(T(B.t.x), T(B.t.y), T(B.t.z));
  const v3<T> bB = V(T(B.b.x), T(B.b.y), T(B.b.z));
  const v3<T> txb = cross(tB, bB);
  const v3<T> ba = V(T(A.b.x), T(A.b.y), T(A.b.z));
  const v3<T> tA = V(T(A.t.x), T(A.t.y), T(A.t.z));
  const T LB = dot(dB, tB);
  const T tba = dot(tB, ba);
  // (sigma b) at the two Gauss points
  const v3<T> sa = seg_stress_times(K, r0 + g0 * dA, LB, tB, bB, txb, tba, ba);
  const v3<T> sc = seg_stress_times(K, r0 + g1 * dA, LB, tB, bB, txb, tba, ba);
  // (sigma b) x t L w, w = 1/2 per node of the two-point rule
  const T hw = T(0.5) * q_rcp(T(A.tN));
  const v3<T> f1 = cross(sa * g1 + sc * g0, tA) * hw;  // weight (1 - u)
  const v3<T> f2 = cross(sa * g0 + sc * g1, tA) * hw;  // weight u
  FA1 = V(double(f1.x), double(f1.y), double(f1.z));
  FA2 = V(double(f2.x), double(f2.y), double(f2.z));
}

}  // namespace ddd
