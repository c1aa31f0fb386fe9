// CPU oracle for the DDD.jl force path.  TEST INFRASTRUCTURE ONLY.
//
// This file restates, in C++, the arithmetic of the reference's force hot path
// so that the HIP kernels can be checked against it.  Only tests/, the smoke
// check and the CPU-baseline leg of the benchmark may load it; the product path
// never does.  Every routine cites the reference line range it follows.
//
// Everything is templated on the scalar type T so that one restatement serves
// four purposes: T=double is the oracle proper; T=long double is the
// conditioning check (how far rounding alone moves a result); T=float is the
// error study for the mixed-precision path; an operation-counting type derives
// the per-pair flop figures used for the roofline.
//
// Every function is pinned by the reference's golden vectors, closed forms or
// invariants (tests/test_oracle*.py).  The closure of calc_uTilde's segments
// with an external node (orc_u_tilde's surface projection,
// DislocationFEM.jl:437-528), for which the reference holds no vectors, is
// pinned by geometry: closed-form surface points, a hand-worked trace of the
// reference's face-plane walk, and an independent solid-angle evaluation of
// the closure polygon, equal up to whole Burgers vectors.
#pragma once
#include <cmath>
#include <cstdint>
#include <cfloat>
#include <limits>

// log and atan of the double-precision oracle.  The reference calls Julia's
// Base.Math log/atan, which no build here can reproduce bit for bit; any
// correctly-specified <1 ulp implementation stands in for them.  The oracle
// uses the fdlibm-method functions that the GPU kernels use, so that both
// sides round identically and agree per pair; those functions are pinned
// against glibc and mpmath in tests/test_oracle.py.  orc_set_libm(1)
// switches the oracle to glibc's log/atan to measure how much the choice of
// implementation moves a result.
#include "../ddd.jl_amd/csrc/ddd_libm.hpp"

namespace ddd_oracle {

extern int g_glibc_libm;
inline double ulog(double x) { return g_glibc_libm ? std::log(x) : ddd_libm::log(x); }
inline double uatan(double x) { return g_glibc_libm ? std::atan(x) : ddd_libm::atan(x); }
inline long double ulog(long double x) { return std::log(x); }
inline long double uatan(long double x) { return std::atan(x); }
inline float ulog(float x) { return std::log(x); }
inline float uatan(float x) { return std::atan(x); }
template <class T> inline T ulog(T x) { return log(x); }
template <class T> inline T uatan(T x) { return atan(x); }

template <class T> struct V3 {
  T x, y, z;
};
template <class T> inline V3<T> mk(T a, T b, T c) { return V3<T>{a, b, c}; }
template <class T> inline V3<T> operator+(const V3<T>& a, const V3<T>& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(const V3<T>& a, const V3<T>& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(const V3<T>& a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(const T& s, const V3<T>& a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator*(const V3<T>& a, const T& s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator/(const V3<T>& a, const T& s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline T dot(const V3<T>& a, const V3<T>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// LinearAlgebra.cross order: (a2 b3 - a3 b2, a3 b1 - a1 b3, a1 b2 - a2 b1).
template <class T> inline V3<T> cross(const V3<T>& a, const V3<T>& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T norm(const V3<T>& a) { using std::sqrt; return sqrt(dot(a, a)); }

// Constants the pair kernel receives (DislocationForces.jl:670-718,
// MaterialConstructor.jl:20-29): aSq, mu/4pi, mu/8pi, aSq*mu/8pi,
// mu/(4pi(1-nu)), aSq*mu/(4pi(1-nu)).
template <class T> struct SegConsts {
  T aSq, mu4pi, mu8pi, mu8piaSq, mu4pinu, mu4pinuaSq;
};

template <class T> struct Forces4 {
  V3<T> f1, f2, f3, f4;
};

// sqrt(eps(Float64)): the branch threshold of the reference, a constant.
template <class T> inline T sqrt_eps() {
  return T(std::sqrt(std::numeric_limits<double>::epsilon()));
}

// Antiderivative values at one corner (x, y): DislocationForces.jl:1480-1561.
template <class T>
inline void seg_seg_integ(T aSq, T d, T c, T cSq, T omcSq, T omcSqI, T x, T y, T* integ) {
  using std::sqrt;
  T aSq_dSq = aSq + d * d * omcSq;
  T xSq = x * x;
  T ySq = y * y;
  T Ra = sqrt(aSq_dSq + xSq + ySq + T(2) * x * y * c);
  T RaInv = T(1) / Ra;

  T Ra_Rd_t1 = Ra + y + x * c;
  T Ra_Rd_t2 = Ra + x + y * c;

  T log_Ra_Rd_t1 = ulog(Ra_Rd_t1);
  T xlog_Ra_Rd_t1 = x * log_Ra_Rd_t1;

  T log_Ra_Rd_t2 = ulog(Ra_Rd_t2);
  T ylog_Ra_Rd_t2 = y * log_Ra_Rd_t2;

  T RaSq_R_t1_I = RaInv / Ra_Rd_t1;
  T xRaSq_R_t1_I = x * RaSq_R_t1_I;
  T xSqRaSq_R_t1_I = x * xRaSq_R_t1_I;

  T RaSq_R_t2_I = RaInv / Ra_Rd_t2;
  T yRaSq_R_t2_I = y * RaSq_R_t2_I;
  T ySqRaSq_R_t2_I = y * yRaSq_R_t2_I;

  T den = T(1) / sqrt(omcSq * aSq_dSq);

  T i1 = T(-2) * den * uatan((T(1) + c) * (Ra + x + y) * den);

  T c_1 = aSq_dSq * i1;
  T c_5_6 = (c * Ra - c_1) * omcSqI;

  T i2 = (c * log_Ra_Rd_t2 - log_Ra_Rd_t1) * omcSqI;
  T i3 = (c * log_Ra_Rd_t1 - log_Ra_Rd_t2) * omcSqI;
  T i4 = (c * c_1 - Ra) * omcSqI;
  T i5 = ylog_Ra_Rd_t2 + c_5_6;
  T i6 = xlog_Ra_Rd_t1 + c_5_6;

  T c_11_12 = i1 - c * RaInv;
  T c_15_18 = c * xRaSq_R_t1_I - RaInv;
  T x_13_14 = x * c_15_18;
  T c_19 = c * yRaSq_R_t2_I - RaInv;
  T y_13_14 = y * c_19;
  T c_16 = log_Ra_Rd_t2 - (x - c * y) * RaInv - cSq * ySqRaSq_R_t2_I;
  T z_15_18 = y * c_16;
  T c_17_19 = log_Ra_Rd_t1 - (y - c * x) * RaInv - cSq * xSqRaSq_R_t1_I;

  T c15_18_19 = T(2) * i4;

  integ[0] = i1;
  integ[1] = i2;
  integ[2] = i3;
  integ[3] = i4;
  integ[4] = i5;
  integ[5] = i6;
  integ[6] = (i1 - xRaSq_R_t1_I - yRaSq_R_t2_I) / (aSq_dSq);
  integ[7] = (RaSq_R_t1_I - c * RaSq_R_t2_I) * omcSqI;
  integ[8] = (RaSq_R_t2_I - c * RaSq_R_t1_I) * omcSqI;
  integ[9] = (RaInv - c * (xRaSq_R_t1_I + yRaSq_R_t2_I + i1)) * omcSqI;
  integ[10] = (xRaSq_R_t1_I + cSq * yRaSq_R_t2_I + c_11_12) * omcSqI;
  integ[11] = (yRaSq_R_t2_I + cSq * xRaSq_R_t1_I + c_11_12) * omcSqI;
  integ[12] = (i3 - x_13_14 + c * (y_13_14 - i2)) * omcSqI;
  integ[13] = (i2 - y_13_14 + c * (x_13_14 - i3)) * omcSqI;
  integ[14] = (i5 - z_15_18 + c * (xSq * c_15_18 - c15_18_19)) * omcSqI;
  integ[15] = (xSqRaSq_R_t1_I + c * c_16 + T(2) * i2) * omcSqI;
  integ[16] = (ySqRaSq_R_t2_I + c * c_17_19 + T(2) * i3) * omcSqI;
  integ[17] = (c15_18_19 - xSq * c_15_18 + c * (z_15_18 - i5)) * omcSqI;
  integ[18] = (c15_18_19 - ySq * c_19 + c * (x * c_17_19 - i6)) * omcSqI;
}

// Near-parallel antiderivatives: DislocationForces.jl:1444-1478.
template <class T>
inline void par_seg_seg_integ(T aSq_dSq, T aSq_dSqI, T x, T y, T* integ) {
  using std::sqrt;
  T xpy = x + y;
  T xmy = x - y;
  T Ra = sqrt(aSq_dSq + xpy * xpy);
  T RaInv = T(1) / Ra;
  T Log_Ra_ypz = ulog(Ra + xpy);
  T tmp = xmy * Ra * aSq_dSqI;
  integ[0] = Ra * aSq_dSqI;
  integ[1] = T(-0.5) * (Log_Ra_ypz - tmp);
  integ[2] = T(-0.5) * (Log_Ra_ypz + tmp);
  integ[3] = -Log_Ra_ypz;
  integ[4] = y * Log_Ra_ypz - Ra;
  integ[5] = x * Log_Ra_ypz - Ra;
  integ[6] = aSq_dSqI * (T(2) * aSq_dSqI * Ra - RaInv);
  integ[7] = aSq_dSqI * (tmp - x * RaInv);
  integ[8] = -aSq_dSqI * (tmp + y * RaInv);
  T i10 = -aSq_dSqI * xpy * RaInv;
  integ[9] = i10;
  integ[10] = RaInv - y * i10;
  integ[11] = RaInv - x * i10;
}

template <class T> inline V3<T> sum11(const V3<T>* V, const T* F) {
  // V7*Fint1 + V8*Fint2 + ... + V17*Fint11, summed left to right.
  V3<T> s = V[0] * F[0];
  for (int k = 1; k < 11; ++k) s = s + V[k] * F[k];
  return s;
}

template <class T>
Forces4<T> seg_seg_force(const SegConsts<T>& K, V3<T> b1, V3<T> n11, V3<T> n12, V3<T> b2,
                         V3<T> n21, V3<T> n22, int depth = 0);

// Generic (non-parallel) pair: DislocationForces.jl:879-1146.
template <class T>
Forces4<T> seg_seg_force_generic(const SegConsts<T>& K, const V3<T>& b1, const V3<T>& n11,
                                 const V3<T>& n12, const V3<T>& b2, const V3<T>& n21,
                                 const V3<T>& n22, T c, T cSq, T omcSq, V3<T> t1, T t1N,
                                 V3<T> t2, T t2N) {
  const T aSq = K.aSq, mu4pi = K.mu4pi, mu8piaSq = K.mu8piaSq, mu4pinu = K.mu4pinu,
          mu4pinuaSq = K.mu4pinuaSq;
  T omcSqI = T(1) / omcSq;

  V3<T> t2ct1 = cross(t2, t1);
  V3<T> t1ct2 = -t2ct1;
  V3<T> b2ct2 = cross(b2, t2);
  V3<T> b1ct1 = cross(b1, t1);

  T t2db2 = dot(t2, b2);
  T t2db1 = dot(t2, b1);
  T t1db2 = dot(t1, b2);
  T t1db1 = dot(t1, b1);

  T t2ct1db2 = dot(t2ct1, b2);
  T t1ct2db1 = dot(t1ct2, b1);
  T b1ct1db2 = dot(b1ct1, b2);
  T b2ct2db1 = dot(b2ct2, b1);

  V3<T> t2ct1ct2 = t1 - c * t2;
  V3<T> t1ct2ct1 = t2 - c * t1;
  V3<T> t2cb1ct2 = b1 - t2db1 * t2;
  V3<T> t1cb2ct1 = b2 - t1db2 * t1;
  V3<T> b1ct1ct2 = t2db1 * t1 - c * b1;
  V3<T> b2ct2ct1 = t1db2 * t2 - c * b2;

  T t2ct1cb1dt1 = t2db1 - t1db1 * c;
  T t1ct2cb2dt2 = t1db2 - t2db2 * c;
  T t2ct1cb1db2 = t2db1 * t1db2 - t1db1 * t2db2;

  V3<T> R1 = n21 - n11;
  V3<T> R2 = n22 - n12;
  T d = dot(R2, t2ct1) * omcSqI;

  T mu4pid = mu4pi * d;
  T mu4pinud = mu4pinu * d;
  T mu4pinudSq = mu4pinud * d;
  T mu4pinudCu = mu4pinudSq * d;
  T mu4pinuaSqd = mu4pinuaSq * d;
  T mu8piaSqd = mu8piaSq * d;

  T lim11 = dot(R1, t1);
  T lim12 = dot(R1, t2);
  T lim21 = dot(R2, t1);
  T lim22 = dot(R2, t2);

  T x1 = (lim12 - c * lim11) * omcSqI;
  T x2 = (lim22 - c * lim21) * omcSqI;
  T y1 = (lim11 - c * lim12) * omcSqI;
  T y2 = (lim21 - c * lim22) * omcSqI;

  T integ[19], tmpi[19];
  seg_seg_integ(aSq, d, c, cSq, omcSq, omcSqI, x1, y1, integ);
  seg_seg_integ(aSq, d, c, cSq, omcSq, omcSqI, x1, y2, tmpi);
  for (int k = 0; k < 19; ++k) integ[k] = integ[k] - tmpi[k];
  seg_seg_integ(aSq, d, c, cSq, omcSq, omcSqI, x2, y1, tmpi);
  for (int k = 0; k < 19; ++k) integ[k] = integ[k] - tmpi[k];
  seg_seg_integ(aSq, d, c, cSq, omcSq, omcSqI, x2, y2, tmpi);
  for (int k = 0; k < 19; ++k) integ[k] = integ[k] + tmpi[k];
  const T* I = integ - 1;  // 1-based view, as in the reference

  Forces4<T> out;
  V3<T> V[11];
  T F[11];
  T tmp1, tmp2;
  V3<T> tmpVec1;
  {  // Seg 1, nodes 2-1: DislocationForces.jl:953-1048
    tmp1 = t1db2 * t2db1 + t2ct1cb1db2;
    V3<T> V1 = tmp1 * t2ct1;
    V3<T> V2 = b1ct1 * t1ct2cb2dt2;
    V3<T> V3_ = t1ct2 * b1ct1db2 - t1cb2ct1 * t2db1;
    V3<T> V4 = -b1ct1 * t2ct1db2;
    V3<T> V5 = b2ct2ct1 * t2db1 - t1ct2 * b2ct2db1;

    tmp1 = mu4pinud * t1ct2db1;
    tmp2 = mu4pinud * b2ct2db1;
    V[0] = mu4pid * V1 - mu4pinud * V2 + tmp1 * b2ct2ct1 + tmp2 * t1ct2ct1;
    tmp1 = mu4pinu * t2db1;
    tmp2 = mu4pinu * b2ct2db1;
    V[1] = mu4pi * V5 - tmp1 * b2ct2ct1 + tmp2 * t1ct2;
    tmp1 = mu4pinu * t1db1;
    V[2] = -tmp1 * b2ct2ct1 + mu4pi * V3_ - mu4pinu * V4;
    tmp1 = mu4pinudCu * t1ct2cb2dt2 * t1ct2db1;
    V[3] = mu8piaSqd * V1 - mu4pinuaSqd * V2 - tmp1 * t1ct2ct1;
    tmp1 = mu4pinudSq * t1ct2cb2dt2 * t2db1;
    tmp2 = mu4pinudSq * t1ct2cb2dt2 * t1ct2db1;
    V[4] = mu8piaSq * V5 + tmp1 * t1ct2ct1 - tmp2 * t1ct2;
    tmp1 = mu4pinudSq * (t2ct1db2 * t1ct2db1 + t1ct2cb2dt2 * t1db1);
    V[5] = mu8piaSq * V3_ - mu4pinuaSq * V4 + tmp1 * t1ct2ct1;
    tmp1 = mu4pinud * (t1ct2cb2dt2 * t1db1 + t2ct1db2 * t1ct2db1);
    tmp2 = mu4pinud * t2ct1db2 * t2db1;
    V[6] = tmp1 * t1ct2 - tmp2 * t1ct2ct1;
    tmp1 = mu4pinud * t1ct2cb2dt2 * t2db1;
    V[7] = tmp1 * t1ct2;
    tmp1 = mu4pinud * t2ct1db2 * t1db1;
    V[8] = -tmp1 * t1ct2ct1;
    tmpVec1 = (mu4pinu * t2ct1db2) * t1ct2;
    V[9] = -tmpVec1 * t2db1;
    V[10] = -tmpVec1 * t1db1;

    F[0] = I[3] - y2 * I[1];
    F[1] = I[4] - y2 * I[2];
    F[2] = I[6] - y2 * I[3];
    F[3] = I[9] - y2 * I[7];
    F[4] = I[10] - y2 * I[8];
    F[5] = I[12] - y2 * I[9];
    F[6] = I[14] - y2 * I[10];
    F[7] = I[13] - y2 * I[11];
    F[8] = I[17] - y2 * I[12];
    F[9] = I[15] - y2 * I[13];
    F[10] = I[19] - y2 * I[14];
    out.f1 = sum11(V, F) * t1N;

    F[0] = y1 * I[1] - I[3];
    F[1] = y1 * I[2] - I[4];
    F[2] = y1 * I[3] - I[6];
    F[3] = y1 * I[7] - I[9];
    F[4] = y1 * I[8] - I[10];
    F[5] = y1 * I[9] - I[12];
    F[6] = y1 * I[10] - I[14];
    F[7] = y1 * I[11] - I[13];
    F[8] = y1 * I[12] - I[17];
    F[9] = y1 * I[13] - I[15];
    F[10] = y1 * I[14] - I[19];
    out.f2 = sum11(V, F) * t1N;
  }
  {  // Seg 2, nodes 4-3: DislocationForces.jl:1050-1145
    tmp1 = t2db1 * t1db2 + t2ct1cb1db2;
    V3<T> V1 = tmp1 * t1ct2;
    V3<T> V2 = b2ct2 * t2ct1cb1dt1;
    V3<T> V3_ = t2ct1 * b1ct1db2 - b1ct1ct2 * t1db2;
    V3<T> V5 = t2cb1ct2 * t1db2 - t2ct1 * b2ct2db1;
    V3<T> V6 = b2ct2 * t1ct2db1;

    tmp1 = mu4pinud * t2ct1db2;
    tmp2 = mu4pinud * b1ct1db2;
    V[0] = mu4pid * V1 - mu4pinud * V2 + tmp1 * b1ct1ct2 + tmp2 * t2ct1ct2;
    tmp1 = mu4pinu * t2db2;
    V[1] = tmp1 * b1ct1ct2 + mu4pi * V5 - mu4pinu * V6;
    tmp1 = mu4pinu * t1db2;
    tmp2 = mu4pinu * b1ct1db2;
    V[2] = mu4pi * V3_ + tmp1 * b1ct1ct2 - tmp2 * t2ct1;
    tmp1 = mu4pinudCu * t2ct1cb1dt1 * t2ct1db2;
    V[3] = mu8piaSqd * V1 - mu4pinuaSqd * V2 - tmp1 * t2ct1ct2;
    tmp1 = mu4pinudSq * (t1ct2db1 * t2ct1db2 + t2ct1cb1dt1 * t2db2);
    V[4] = mu8piaSq * V5 - mu4pinuaSq * V6 - tmp1 * t2ct1ct2;
    tmp1 = mu4pinudSq * t2ct1cb1dt1 * t1db2;
    tmp2 = mu4pinudSq * t2ct1cb1dt1 * t2ct1db2;
    V[5] = mu8piaSq * V3_ - tmp1 * t2ct1ct2 + tmp2 * t2ct1;
    tmp1 = mu4pinud * (t2ct1cb1dt1 * t2db2 + t1ct2db1 * t2ct1db2);
    tmp2 = mu4pinud * t1ct2db1 * t1db2;
    V[6] = tmp1 * t2ct1 - tmp2 * t2ct1ct2;
    tmp1 = mu4pinud * t1ct2db1 * t2db2;
    V[7] = -tmp1 * t2ct1ct2;
    tmp1 = mu4pinud * t2ct1cb1dt1 * t1db2;
    V[8] = tmp1 * t2ct1;
    tmpVec1 = (mu4pinu * t1ct2db1) * t2ct1;
    V[9] = tmpVec1 * t2db2;
    V[10] = tmpVec1 * t1db2;

    F[0] = x2 * I[1] - I[2];
    F[1] = x2 * I[2] - I[5];
    F[2] = x2 * I[3] - I[4];
    F[3] = x2 * I[7] - I[8];
    F[4] = x2 * I[8] - I[11];
    F[5] = x2 * I[9] - I[10];
    F[6] = x2 * I[10] - I[13];
    F[7] = x2 * I[11] - I[16];
    F[8] = x2 * I[12] - I[14];
    F[9] = x2 * I[13] - I[18];
    F[10] = x2 * I[14] - I[15];
    out.f3 = sum11(V, F) * t2N;

    F[0] = I[2] - x1 * I[1];
    F[1] = I[5] - x1 * I[2];
    F[2] = I[4] - x1 * I[3];
    F[3] = I[8] - x1 * I[7];
    F[4] = I[11] - x1 * I[8];
    F[5] = I[10] - x1 * I[9];
    F[6] = I[13] - x1 * I[10];
    F[7] = I[16] - x1 * I[11];
    F[8] = I[14] - x1 * I[12];
    F[9] = I[18] - x1 * I[13];
    F[10] = I[15] - x1 * I[14];
    out.f4 = sum11(V, F) * t2N;
  }
  (void)n11; (void)n21; (void)b1; (void)b2;
  return out;
}

// Near-parallel pair: DislocationForces.jl:1167-1442.  The reference's
// recursion into the generic routine for the tilted correction segments
// (:1284-1317, :1400-1433) is kept as a recursive call.
template <class T>
Forces4<T> seg_seg_force_parallel(const SegConsts<T>& K, V3<T> b1, V3<T> n11, V3<T> n12,
                                  V3<T> b2, V3<T> n21, V3<T> n22, int depth) {
  using std::sqrt;
  const T mu4pi = K.mu4pi, mu8piaSq = K.mu8piaSq, mu4pinu = K.mu4pinu,
          mu4pinuaSq = K.mu4pinuaSq, aSq = K.aSq;
  bool flip = false;
  V3<T> t2 = n22 - n21;
  T t2N = T(1) / norm(t2);
  t2 = t2 * t2N;
  V3<T> t1 = n12 - n11;
  T t1N = T(1) / norm(t1);
  t1 = t1 * t1N;
  T c = dot(t2, t1);
  const T se = sqrt_eps<T>();
  T hCotanThc = sqrt((T(1) - se * T(1.01)) / (se * T(1.01))) / T(2);
  if (c < T(0)) {
    flip = true;
    V3<T> tmp = n12; n12 = n11; n11 = tmp;
    t1 = -t1;
    b1 = -b1;
  }
  Forces4<T> out;
  T integ[12], tmpi[12];
  {  // Segment 1 side: :1204-1318
    T tmp = dot(n22 - n21, t1);
    V3<T> n22m = n21 + tmp * t1;
    V3<T> diff = n22 - n22m;
    T magDiff = norm(diff);
    V3<T> tmpVec1 = T(0.5) * diff;
    V3<T> tmpVec2 = (hCotanThc * magDiff) * t1;
    V3<T> n21m = n21 + tmpVec1 + tmpVec2;
    n22m = n22m + tmpVec1 - tmpVec2;

    V3<T> R = n21m - n11;
    T Rdt1 = dot(R, t1);
    V3<T> nd = R - Rdt1 * t1;
    T ndb1 = dot(nd, b1); (void)ndb1;
    T dSq = dot(nd, nd);
    T aSq_dSq = aSq + dSq;
    T aSq_dSqI = T(1) / aSq_dSq;
    T x1 = dot(n21m, t1);
    T x2 = dot(n22m, t1);
    T y1 = -dot(n11, t1);
    T y2 = -dot(n12, t1);
    T t1db2 = dot(t1, b2);
    T t1db1 = dot(t1, b1);
    T nddb1 = dot(nd, b1);
    V3<T> b2ct1 = cross(b2, t1);
    V3<T> b1ct1 = cross(b1, t1);
    V3<T> ndct1 = cross(nd, t1);
    T b2ct1db1 = dot(b2ct1, b1);
    T b2ct1dnd = dot(b2ct1, nd);
    V3<T> b2ct1ct1 = t1db2 * t1 - b2;

    par_seg_seg_integ(aSq_dSq, aSq_dSqI, x1, y1, integ);
    par_seg_seg_integ(aSq_dSq, aSq_dSqI, x1, y2, tmpi);
    for (int k = 0; k < 12; ++k) integ[k] = integ[k] - tmpi[k];
    par_seg_seg_integ(aSq_dSq, aSq_dSqI, x2, y1, tmpi);
    for (int k = 0; k < 12; ++k) integ[k] = integ[k] - tmpi[k];
    par_seg_seg_integ(aSq_dSq, aSq_dSqI, x2, y2, tmpi);
    for (int k = 0; k < 12; ++k) integ[k] = integ[k] + tmpi[k];
    const T* I = integ - 1;

    T tq = t1db1 * t1db2;
    V3<T> tv1 = tq * nd;
    V3<T> tv2 = b2ct1dnd * b1ct1;
    V3<T> V1 = mu4pinu * (nddb1 * b2ct1ct1 + b2ct1db1 * ndct1 - tv2) - mu4pi * tv1;
    tq = (mu4pinu - mu4pi) * t1db1;
    V3<T> V2 = tq * b2ct1ct1;
    tq = mu4pinu * b2ct1dnd * nddb1;
    V3<T> V3_ = -mu8piaSq * tv1 - mu4pinuaSq * tv2 - tq * ndct1;
    tq = mu8piaSq * t1db1;
    T tq2 = mu4pinu * b2ct1dnd * t1db1;
    V3<T> V4 = -tq * b2ct1ct1 - tq2 * ndct1;

    T F1 = I[3] - y1 * I[1], F2 = I[6] - y1 * I[4], F3 = I[9] - y1 * I[7], F4 = I[12] - y1 * I[10];
    out.f2 = (V1 * F1 + V2 * F2 + V3_ * F3 + V4 * F4) * t1N;
    F1 = y2 * I[1] - I[3]; F2 = y2 * I[4] - I[6]; F3 = y2 * I[7] - I[9]; F4 = y2 * I[10] - I[12];
    out.f1 = (V1 * F1 + V2 * F2 + V3_ * F3 + V4 * F4) * t1N;

    T magDiffSq = dot(diff, diff);
    T magn21mSq = dot(n21m, n21m);
    T magn22mSq = dot(n22m, n22m);
    if (magDiffSq > se * (magn21mSq + magn22mSq)) {
      Forces4<T> s = seg_seg_force(K, b2, n21, n21m, b1, n11, n12, depth + 1);
      out.f1 = out.f1 + s.f3;
      out.f2 = out.f2 + s.f4;
      s = seg_seg_force(K, b2, n22m, n22, b1, n11, n12, depth + 1);
      out.f1 = out.f1 + s.f3;
      out.f2 = out.f2 + s.f4;
    }
  }
  {  // Segment 2 side: :1320-1434
    T tmp = dot(n12 - n11, t2);
    V3<T> n12m = n11 + tmp * t2;
    V3<T> diff = n12 - n12m;
    T magDiff = norm(diff);
    V3<T> tmpVec1 = T(0.5) * diff;
    V3<T> tmpVec2 = (hCotanThc * magDiff) * t2;
    V3<T> n11m = n11 + tmpVec1 + tmpVec2;
    n12m = n12m + tmpVec1 - tmpVec2;

    V3<T> R = n21 - n11m;
    T Rdt2 = dot(R, t2);
    V3<T> nd = R - Rdt2 * t2;
    T dSq = dot(nd, nd);
    T aSq_dSq = aSq + dSq;
    T aSq_dSqI = T(1) / aSq_dSq;
    T x1 = dot(n21, t2);
    T x2 = dot(n22, t2);
    T y1 = -dot(n11m, t2);
    T y2 = -dot(n12m, t2);
    T t2db2 = dot(t2, b2);
    T t2db1 = dot(t2, b1);
    T nddb2 = dot(nd, b2);
    V3<T> b2ct2 = cross(b2, t2);
    V3<T> b1ct2 = cross(b1, t2);
    V3<T> ndct2 = cross(nd, t2);
    T b1ct2db2 = dot(b1ct2, b2);
    T b1ct2dnd = dot(b1ct2, nd);
    V3<T> b1ct2ct2 = t2db1 * t2 - b1;

    par_seg_seg_integ(aSq_dSq, aSq_dSqI, x1, y1, integ);
    par_seg_seg_integ(aSq_dSq, aSq_dSqI, x1, y2, tmpi);
    for (int k = 0; k < 12; ++k) integ[k] = integ[k] - tmpi[k];
    par_seg_seg_integ(aSq_dSq, aSq_dSqI, x2, y1, tmpi);
    for (int k = 0; k < 12; ++k) integ[k] = integ[k] - tmpi[k];
    par_seg_seg_integ(aSq_dSq, aSq_dSqI, x2, y2, tmpi);
    for (int k = 0; k < 12; ++k) integ[k] = integ[k] + tmpi[k];
    const T* I = integ - 1;

    T tq = t2db2 * t2db1;
    V3<T> tv1 = tq * nd;
    V3<T> tv2 = b1ct2dnd * b2ct2;
    V3<T> V1 = mu4pinu * (nddb2 * b1ct2ct2 + b1ct2db2 * ndct2 - tv2) - mu4pi * tv1;
    tq = (mu4pinu - mu4pi) * t2db2;
    V3<T> V2 = tq * b1ct2ct2;
    tq = mu4pinu * b1ct2dnd * nddb2;
    V3<T> V3_ = -mu8piaSq * tv1 - mu4pinuaSq * tv2 - tq * ndct2;
    tq = mu8piaSq * t2db2;
    T tq2 = mu4pinu * b1ct2dnd * t2db2;
    V3<T> V4 = -tq * b1ct2ct2 - tq2 * ndct2;

    T F1 = I[2] - x1 * I[1], F2 = I[5] - x1 * I[4], F3 = I[8] - x1 * I[7], F4 = I[11] - x1 * I[10];
    out.f4 = (V1 * F1 + V2 * F2 + V3_ * F3 + V4 * F4) * t2N;
    F1 = x2 * I[1] - I[2]; F2 = x2 * I[4] - I[5]; F3 = x2 * I[7] - I[8]; F4 = x2 * I[10] - I[11];
    out.f3 = (V1 * F1 + V2 * F2 + V3_ * F3 + V4 * F4) * t2N;

    T magDiffSq = magDiff * magDiff;
    T magn11mSq = dot(n11m, n11m);
    T magn12mSq = dot(n12m, n12m);
    if (magDiffSq > se * (magn11mSq + magn12mSq)) {
      Forces4<T> s = seg_seg_force(K, b1, n11, n11m, b2, n21, n22, depth + 1);
      out.f3 = out.f3 + s.f3;
      out.f4 = out.f4 + s.f4;
      s = seg_seg_force(K, b1, n12m, n12, b2, n21, n22, depth + 1);
      out.f3 = out.f3 + s.f3;
      out.f4 = out.f4 + s.f4;
    }
  }
  if (flip) {
    V3<T> tmp = out.f1; out.f1 = out.f2; out.f2 = tmp;
  }
  return out;
}

// Dispatcher: DislocationForces.jl:879-892 (normalise, test 1-c^2 > sqrt(eps)).
template <class T>
Forces4<T> seg_seg_force(const SegConsts<T>& K, V3<T> b1, V3<T> n11, V3<T> n12, V3<T> b2,
                         V3<T> n21, V3<T> n22, int depth) {
  V3<T> t2 = n22 - n21;
  T t2N = T(1) / norm(t2);
  t2 = t2 * t2N;
  V3<T> t1 = n12 - n11;
  T t1N = T(1) / norm(t1);
  t1 = t1 * t1N;
  T c = dot(t1, t2);
  T cSq = c * c;
  T omcSq = T(1) - cSq;
  if (omcSq > sqrt_eps<T>()) {
    return seg_seg_force_generic(K, b1, n11, n12, b2, n21, n22, c, cSq, omcSq, t1, t1N, t2, t2N);
  }
  return seg_seg_force_parallel(K, b1, n11, n12, b2, n21, n22, depth);
}

// True when the reference would take the near-parallel branch for this pair.
template <class T>
bool seg_seg_is_parallel(V3<T> n11, V3<T> n12, V3<T> n21, V3<T> n22) {
  V3<T> t2 = n22 - n21;
  t2 = t2 * (T(1) / norm(t2));
  V3<T> t1 = n12 - n11;
  t1 = t1 * (T(1) / norm(t1));
  T c = dot(t1, t2);
  return !(T(1) - c * c > sqrt_eps<T>());
}

// Self force of one segment: DislocationForces.jl:414-455.  Returns the
// force on node 2; node 1 receives its negative.
template <class T>
V3<T> self_force(V3<T> tVec, V3<T> b, T a, T aSq, T mu4pi, T nuOmNuInv, T Ec, T omNuInv) {
  using std::sqrt;
  T tVecSq = dot(tVec, tVec);
  T L = sqrt(tVecSq);
  T Linv = T(1) / L;
  V3<T> t = tVec * Linv;
  T La = sqrt(tVecSq + aSq);
  T bScrew = dot(t, b);
  V3<T> bEdge = b - bScrew * t;
  T bEdgeSq = dot(bEdge, bEdge);
  T LaMa = La - a;
  T tor = mu4pi * bScrew *
          (nuOmNuInv * (ulog((La + L) / a) - T(2) * LaMa * Linv) - LaMa * LaMa / (T(2) * La * L));
  T torCore = T(2) * Ec * nuOmNuInv * bScrew;
  T torTot = tor + torCore;
  T lonCore = (bScrew * bScrew + bEdgeSq * omNuInv) * Ec;
  return torTot * bEdge - lonCore * t;
}

// Non-singular stress of one straight segment at one point, added into
// sig[6] in the order (xx, yy, zz, xy, xz, yz): DislocationFEM.jl:129-299.
template <class T>
void sigma_tilde_point(const V3<T>& b, const V3<T>& n1, const V3<T>& n2, const V3<T>& x,
                       T a2, T mu4pi, T mu4pinu, T nuMu4pinu, T a2mu8pi, T* sig) {
  using std::sqrt;
  V3<T> tv = n2 - n1;
  V3<T> t = tv / norm(tv);
  V3<T> R = x - n1;
  T Rdt = dot(R, t);
  V3<T> nd = R - Rdt * t;
  T d2 = dot(nd, nd);
  T s1 = -Rdt;
  T s2 = -dot(x - n2, t);
  T a2_d2 = a2 + d2;
  T a2d2inv = T(1) / a2_d2;

  T Ra = sqrt(a2_d2 + s1 * s1);
  T Rainv = T(1) / Ra;
  T Ra3inv = Rainv * Rainv * Rainv;
  T sRa3inv = s1 * Ra3inv;
  T s_03a = s1 * Rainv * a2d2inv;
  T s_13a = -Rainv;
  T s_05a = (T(2) * s_03a + sRa3inv) * a2d2inv;
  T s_15a = -Ra3inv;
  T s_25a = s_03a - sRa3inv;

  Ra = sqrt(a2_d2 + s2 * s2);
  Rainv = T(1) / Ra;
  Ra3inv = Rainv * Rainv * Rainv;
  sRa3inv = s2 * Ra3inv;
  T s_03b = s2 * Rainv * a2d2inv;
  T s_13b = -Rainv;
  T s_05b = (T(2) * s_03b + sRa3inv) * a2d2inv;
  T s_15b = -Ra3inv;
  T s_25b = s_03b - sRa3inv;

  T s_03 = s_03b - s_03a;
  T s_13 = s_13b - s_13a;
  T s_05 = s_05b - s_05a;
  T s_15 = s_15b - s_15a;
  T s_25 = s_25b - s_25a;

  V3<T> txb = cross(t, b);
  V3<T> dxb = cross(nd, b);
  T dxbdt = dot(dxb, t);
  const T tt[3] = {t.x, t.y, t.z};
  const T dd[3] = {nd.x, nd.y, nd.z};
  const T xb[3] = {txb.x, txb.y, txb.z};
  const T db[3] = {dxb.x, dxb.y, dxb.z};
  // Component pairs in output order xx, yy, zz, xy, xz, yz.
  const int P[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
  T common = mu4pinu * dxbdt;
  for (int c = 0; c < 6; ++c) {
    int p = P[c][0], q = P[c][1];
    bool diag = (p == q);
    T dmd = dd[p] * dd[q];
    T tmt = tt[p] * tt[q];
    T tmd = diag ? T(2) * tt[p] * dd[p] : tt[p] * dd[q] + tt[q] * dd[p];
    T tmtxb = diag ? T(2) * tt[p] * xb[p] : tt[p] * xb[q] + tt[q] * xb[p];
    T dmtxb = diag ? T(2) * dd[p] * xb[p] : dd[p] * xb[q] + dd[q] * xb[p];
    T tmdxb = diag ? T(2) * tt[p] * db[p] : tt[p] * db[q] + tt[q] * db[p];
    T I03 = diag ? common + mu4pinu * dmtxb - mu4pi * tmdxb : mu4pinu * dmtxb - mu4pi * tmdxb;
    T I13 = -nuMu4pinu * tmtxb;
    T I05 = diag ? common * (a2 + dmd) - a2mu8pi * tmdxb : common * dmd - a2mu8pi * tmdxb;
    T I15 = a2mu8pi * tmtxb - common * tmd;
    T I25 = common * tmt;
    sig[c] = sig[c] + (I03 * s_03 + I13 * s_13 + I05 * s_05 + I15 * s_15 + I25 * s_25);
  }
}

// Far-pair force by quadrature (the mixed-precision path's formulation, no
// reference counterpart): the nodal forces on segment A from segment B are
// F_A1 = L_A * sum_q w_q (1 - u_q) (sigma_B(x_A(u_q)) b_A) x t_A and
// F_A2 = L_A * sum_q w_q u_q (...), with nq-point Gauss-Legendre nodes u_q on
// [0, 1] and sigma_B the non-singular stress of segment B (sigma_tilde_point).
// Positions enter as differences taken in double, so T = float only sees
// lengths of the pair's own scale.
template <class T> inline void gl_nodes(int nq, T* u, T* w) {
  if (nq == 1) { u[0] = T(0.5); w[0] = T(1); return; }
  if (nq == 2) {
    const double d = 0.5 / std::sqrt(3.0);
    u[0] = T(0.5 - d); u[1] = T(0.5 + d); w[0] = w[1] = T(0.5); return;
  }
  if (nq == 3) {
    const double d = 0.5 * std::sqrt(0.6);
    u[0] = T(0.5 - d); u[1] = T(0.5); u[2] = T(0.5 + d);
    w[0] = w[2] = T(5.0 / 18.0); w[1] = T(8.0 / 18.0); return;
  }
  // nq == 4
  const double a = std::sqrt(3.0 / 7.0 - 2.0 / 7.0 * std::sqrt(6.0 / 5.0));
  const double b = std::sqrt(3.0 / 7.0 + 2.0 / 7.0 * std::sqrt(6.0 / 5.0));
  const double wa = (18.0 + std::sqrt(30.0)) / 36.0, wb = (18.0 - std::sqrt(30.0)) / 36.0;
  u[0] = T(0.5 - 0.5 * b); u[1] = T(0.5 - 0.5 * a); u[2] = T(0.5 + 0.5 * a); u[3] = T(0.5 + 0.5 * b);
  w[0] = w[3] = T(0.5 * wb); w[1] = w[2] = T(0.5 * wa);
}

template <class T>
inline void far_side_quad(const V3<double>& bA, const V3<double>& a1, const V3<double>& a2,
                          const V3<double>& bB, const V3<double>& p1, const V3<double>& p2,
                          int nq, T a2c, T mu4pi, T mu4pinu, T nuMu4pinu, T a2mu8pi,
                          V3<double>& FA1, V3<double>& FA2) {
  T u[4], w[4];
  gl_nodes<T>(nq, u, w);
  const V3<double> dA = a2 - a1, d0 = a1 - p1, dB = p2 - p1;
  const V3<T> tA = mk(T(dA.x), T(dA.y), T(dA.z));
  const T LA = norm(tA);
  const V3<T> tAu = tA / LA;
  const V3<T> r0 = mk(T(d0.x), T(d0.y), T(d0.z));
  const V3<T> q2 = mk(T(dB.x), T(dB.y), T(dB.z));
  const V3<T> b = mk(T(bB.x), T(bB.y), T(bB.z));
  const V3<T> ba = mk(T(bA.x), T(bA.y), T(bA.z));
  const V3<T> zero = mk(T(0), T(0), T(0));
  V3<T> f1 = zero, f2 = zero;
  for (int q = 0; q < nq; ++q) {
    V3<T> x = r0 + u[q] * tA;  // point relative to p1
    T sg[6] = {0, 0, 0, 0, 0, 0};
    sigma_tilde_point(b, zero, q2, x, a2c, mu4pi, mu4pinu, nuMu4pinu, a2mu8pi, sg);
    V3<T> sb = mk(sg[0] * ba.x + sg[3] * ba.y + sg[4] * ba.z, sg[3] * ba.x + sg[1] * ba.y + sg[5] * ba.z,
                  sg[4] * ba.x + sg[5] * ba.y + sg[2] * ba.z);
    V3<T> f = cross(sb, tAu) * (LA * w[q]);
    f1 = f1 + f * (T(1) - u[q]);
    f2 = f2 + f * u[q];
  }
  FA1 = mk(double(f1.x), double(f1.y), double(f1.z));
  FA2 = mk(double(f2.x), double(f2.y), double(f2.z));
}

}  // namespace ddd_oracle
