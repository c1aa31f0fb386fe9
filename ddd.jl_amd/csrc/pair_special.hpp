// The near-parallel branch of the pair force and the segment-record reader,
// shared by the all-pairs tiles and the hybrid near field.
#pragma once
#include "kernels.hpp"

namespace ddd {

static __device__ __forceinline__ Seg<double> seg_from(const double* __restrict__ seg, int64_t ld,
                                                       int64_t i) {
  Seg<double> r;
  r.n1 = V(seg[F_N1X * ld + i], seg[F_N1Y * ld + i], seg[F_N1Z * ld + i]);
  r.n2 = V(seg[F_N2X * ld + i], seg[F_N2Y * ld + i], seg[F_N2Z * ld + i]);
  r.b = V(seg[F_BX * ld + i], seg[F_BY * ld + i], seg[F_BZ * ld + i]);
  r.t = V(seg[F_TX * ld + i], seg[F_TY * ld + i], seg[F_TZ * ld + i]);
  r.tN = seg[F_TN * ld + i];
  return r;
}

// A corrective sub-pair of the near-parallel branch: generic force on the
// nodes of `other` from the straight piece p -> q carrying Burgers vector bp.
// In the reference this is a full calcSegSegForce call whose third and fourth
// outputs are kept (:1285-1317, :1401-1433).  The tilt construction makes the
// piece sit at the critical angle, sin^2 = 1.01 sqrt(eps), so it always takes
// the generic branch; the flag reports it if rounding ever says otherwise.
[[maybe_unused]] static __device__ __noinline__ void corrective_pair(const SegK<double>& K, v3<double> bp, v3<double> p,
                                             v3<double> q, const Seg<double>& other,
                                             v3<double>& Fo1, v3<double>& Fo2,
                                             unsigned* flags) {
  Seg<double> sp;
  sp.n1 = p;
  sp.n2 = q;
  sp.b = bp;
  unit_tangent(p, q, sp.t, sp.tN);
  double c;
  double omcSq = one_minus_c2(sp.t, other.t, c);
  if (!(omcSq > DDD_SQRT_EPS)) atomicOr(flags, ERRBIT_NESTED_PARALLEL);
  v3<double> d1, d2, F3, F4;
  pair_generic<double, false, true>(K, sp, other, c, omcSq, d1, d2, F3, F4);
  Fo1 = Fo1 + F3;
  Fo2 = Fo2 + F4;
}

// calcParSegSegForce (:1167-1442) on device.
[[maybe_unused]] static __device__ __noinline__ void special_pair_dev(const SegK<double>& K, const Seg<double>& s1in,
                                          const Seg<double>& s2, v3<double>& F1, v3<double>& F2,
                                          v3<double>& F3, v3<double>& F4, unsigned* flags) {
  const double c = dot(s2.t, s1in.t);
  const bool flip = c < 0.0;
  Seg<double> s1 = s1in;
  if (flip) {  // :1197-1202; the flipped tangent is the exact negation
    s1.n1 = s1in.n2;
    s1.n2 = s1in.n1;
    s1.t = -s1in.t;
    s1.b = -s1in.b;
  }
  v3<double> c1a, c1b, c2a, c2b, d1a, d1b, d2a, d2b;
  bool corr1, corr2;
  par_side<double>(K, s1.n1, s1.n2, s1.t, s1.tN, s1.b, s2.n1, s2.n2, s2.b, true, F1, F2, c1a,
                   c1b, c2a, c2b, corr1);
  par_side<double>(K, s2.n1, s2.n2, s2.t, s2.tN, s2.b, s1.n1, s1.n2, s1.b, false, F3, F4, d1a,
                   d1b, d2a, d2b, corr2);
  if (corr1) {
    corrective_pair(K, s2.b, c1a, c1b, s1, F1, F2, flags);
    corrective_pair(K, s2.b, c2a, c2b, s1, F1, F2, flags);
  }
  if (corr2) {
    corrective_pair(K, s1.b, d1a, d1b, s2, F3, F4, flags);
    corrective_pair(K, s1.b, d2a, d2b, s2, F3, F4, flags);
  }
  if (flip) {
    v3<double> tmp = F1;
    F1 = F2;
    F2 = tmp;
  }
}


}  // namespace ddd
