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

// The near-parallel branch, calcParSegSegForce (:1167-1442), in two call-free
// passes.  Pass A (special_pair_base) flips segment 1 when c < 0 (:1197-1202),
// evaluates both sides' integrals (par_side), un-flips (:1437-1439) and says
// whether a side needs its corrective generic calls (:1284, :1400).  Pass B
// (special_pair_fix), for the few pairs that do, adds the corrective
// sub-pairs -- full generic pairs of a tilted piece against the other
// segment, of which the reference keeps the third and fourth outputs
// (:1285-1317, :1401-1433) -- to the forces of pass A, in the reference's
// order.  The tilt construction puts each piece at the critical angle,
// sin^2 = 1.01 sqrt(eps), so it takes the generic branch; the flag reports
// it if rounding ever says otherwise.
struct SpecialFrame {
  Seg<double> s1;
  bool flip;
};
static __device__ __forceinline__ SpecialFrame special_frame(const Seg<double>& s1in,
                                                             const Seg<double>& s2) {
  SpecialFrame f{s1in, dot(s2.t, s1in.t) < 0.0};
  if (f.flip) {  // the flipped tangent is the exact negation
    f.s1.n1 = s1in.n2;
    f.s1.n2 = s1in.n1;
    f.s1.t = -s1in.t;
    f.s1.b = -s1in.b;
  }
  return f;
}

// Pass A: F1, F2 on segment 1's nodes and F3, F4 on segment 2's, without the
// corrective calls; returns true when a side needs them.
static __device__ __forceinline__ bool special_pair_base(const SegK<double>& K,
                                                         const Seg<double>& s1in,
                                                         const Seg<double>& s2, v3<double>& F1,
                                                         v3<double>& F2, v3<double>& F3,
                                                         v3<double>& F4) {
  const SpecialFrame fr = special_frame(s1in, s2);
  const Seg<double>& s1 = fr.s1;
  v3<double> c1a, c1b, c2a, c2b, d1a, d1b, d2a, d2b;
  bool corr1, corr2;
  par_side<double>(K, s1.n1, s1.n2, s1.t, s1.tN, s1.b, s2.n1, s2.n2, s2.b, true, F1, F2, c1a,
                   c1b, c2a, c2b, corr1);
  par_side<double>(K, s2.n1, s2.n2, s2.t, s2.tN, s2.b, s1.n1, s1.n2, s1.b, false, F3, F4, d1a,
                   d1b, d2a, d2b, corr2);
  if (fr.flip) {
    const v3<double> tmp = F1;
    F1 = F2;
    F2 = tmp;
  }
  return corr1 || corr2;
}

// Generic force on the nodes of `other` from the straight piece p -> q
// carrying Burgers vector bp, added to (Fo1, Fo2).
static __device__ __forceinline__ void corrective_add(const SegK<double>& K, v3<double> bp,
                                                      v3<double> p, v3<double> q,
                                                      const Seg<double>& other, v3<double>& Fo1,
                                                      v3<double>& Fo2, unsigned* flags) {
  Seg<double> sp;
  sp.n1 = p;
  sp.n2 = q;
  sp.b = bp;
  unit_tangent(p, q, sp.t, sp.tN);
  double c;
  const double omcSq = one_minus_c2(sp.t, other.t, c);
  if (!(omcSq > DDD_SQRT_EPS)) atomicOr(flags, ERRBIT_NESTED_PARALLEL);
  v3<double> d1, d2, F3, F4;
  pair_generic<double, false, true>(K, sp, other, c, omcSq, d1, d2, F3, F4);
  Fo1 = Fo1 + F3;
  Fo2 = Fo2 + F4;
}

// Pass B: the corrective sub-pairs of both sides, added to the forces of
// pass A (segment 1's two nodes swap places when it was flipped).  The
// pieces are p1 -> p1m and p2m -> p2 of the tilt of the other segment onto
// each side's line; the sub-pairs run one at a time.
static __device__ __forceinline__ void special_pair_fix(const SegK<double>& K,
                                                        const Seg<double>& s1in,
                                                        const Seg<double>& s2, v3<double>& F1,
                                                        v3<double>& F2, v3<double>& F3,
                                                        v3<double>& F4, unsigned* flags) {
  const SpecialFrame fr = special_frame(s1in, s2);
  const Seg<double>& s1 = fr.s1;
  const double se = DDD_SQRT_EPS;
  v3<double>& G1 = fr.flip ? F2 : F1;  // the flipped frame's node 1
  v3<double>& G2 = fr.flip ? F1 : F2;
  v3<double> p1m, p2m;
  double m2, m;
  par_tilt(s1.t, s2.n1, s2.n2, p1m, p2m, m2, m);
  if (m2 > se * (dot(p1m, p1m) + dot(p2m, p2m))) {
    corrective_add(K, s2.b, s2.n1, p1m, s1, G1, G2, flags);
    corrective_add(K, s2.b, p2m, s2.n2, s1, G1, G2, flags);
  }
  par_tilt(s2.t, s1.n1, s1.n2, p1m, p2m, m2, m);
  if (m * m > se * (dot(p1m, p1m) + dot(p2m, p2m))) {
    corrective_add(K, s1.b, s1.n1, p1m, s2, F3, F4, flags);
    corrective_add(K, s1.b, p2m, s1.n2, s2, F3, F4, flags);
  }
}

// Both passes for one pair, as one call (the rows-only kernel, which
// evaluates few pairs, keeps it out of line to bound its registers).
[[maybe_unused]] static __device__ __noinline__ void special_pair_dev(const SegK<double>& K,
                                                                      const Seg<double>& s1in,
                                                                      const Seg<double>& s2,
                                                                      v3<double>& F1,
                                                                      v3<double>& F2,
                                                                      v3<double>& F3,
                                                                      v3<double>& F4,
                                                                      unsigned* flags) {
  if (special_pair_base(K, s1in, s2, F1, F2, F3, F4))
    special_pair_fix(K, s1in, s2, F1, F2, F3, F4, flags);
}

}  // namespace ddd
