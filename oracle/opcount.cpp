// Operation counts of the oracle's restatement.  TEST INFRASTRUCTURE ONLY.
//
// Runs the same templated code as the oracle with a scalar type that counts
// every arithmetic operation it performs, so the per-unit work figures used for
// the roofline (flops per generic pair, per self-force segment, per
// segment-point stress evaluation) are derived from the code rather than by
// hand.  add/sub/mul count one flop each; div, sqrt, log and atan are counted
// separately; negation and comparisons are free.
#include "ddd_oracle.hpp"

namespace opc {
struct Counts {
  long add = 0, mul = 0, div = 0, sqrt = 0, log = 0, atan = 0;
};
static Counts g;

struct OC {
  double v;
  OC() : v(0) {}
  OC(double x) : v(x) {}
  explicit operator double() const { return v; }
};
inline OC operator+(OC a, OC b) { ++g.add; return OC(a.v + b.v); }
inline OC operator-(OC a, OC b) { ++g.add; return OC(a.v - b.v); }
inline OC operator*(OC a, OC b) { ++g.mul; return OC(a.v * b.v); }
inline OC operator/(OC a, OC b) { ++g.div; return OC(a.v / b.v); }
inline OC operator-(OC a) { return OC(-a.v); }
inline OC& operator+=(OC& a, OC b) { a = a + b; return a; }
inline bool operator>(OC a, OC b) { return a.v > b.v; }
inline bool operator<(OC a, OC b) { return a.v < b.v; }
inline OC sqrt(OC a) { ++g.sqrt; return OC(std::sqrt(a.v)); }
inline OC log(OC a) { ++g.log; return OC(std::log(a.v)); }
inline OC atan(OC a) { ++g.atan; return OC(std::atan(a.v)); }
}  // namespace opc

using namespace ddd_oracle;
using opc::OC;

static void emit(double* out) {
  out[0] = double(opc::g.add + opc::g.mul);
  out[1] = double(opc::g.add);
  out[2] = double(opc::g.mul);
  out[3] = double(opc::g.div);
  out[4] = double(opc::g.sqrt);
  out[5] = double(opc::g.log);
  out[6] = double(opc::g.atan);
}

extern "C" {

// counts[7] = {add+mul, add, mul, div, sqrt, log, atan} for one pair given by
// the same arguments as orc_segseg_pair (double data).  Returns 1 when the
// near-parallel branch was taken.
int orc_count_segseg_pair(const double* K, const double* b1, const double* n11, const double* n12,
                          const double* b2, const double* n21, const double* n22, double* counts) {
  opc::g = opc::Counts{};
  SegConsts<OC> KK{K[0], K[1], K[2], K[3], K[4], K[5]};
  auto v = [](const double* p) { return mk(OC(p[0]), OC(p[1]), OC(p[2])); };
  Forces4<OC> f = seg_seg_force(KK, v(b1), v(n11), v(n12), v(b2), v(n21), v(n22));
  (void)f;
  emit(counts);
  opc::g = opc::Counts{};
  bool par = seg_seg_is_parallel(v(n11), v(n12), v(n21), v(n22));
  return par ? 1 : 0;
}

// Counts for one segment-point stress evaluation (sigma_tilde_point).
void orc_count_sigma_tilde(const double* sc, const double* b, const double* n1, const double* n2,
                           const double* x, double* counts) {
  opc::g = opc::Counts{};
  auto v = [](const double* p) { return mk(OC(p[0]), OC(p[1]), OC(p[2])); };
  OC sig[6];
  sigma_tilde_point(v(b), v(n1), v(n2), v(x), OC(sc[0]), OC(sc[1]), OC(sc[2]), OC(sc[3]),
                    OC(sc[4]), sig);
  emit(counts);
}

// Counts for one self-force evaluation.
void orc_count_self_force(const double* sc, const double* tv, const double* b, double* counts) {
  opc::g = opc::Counts{};
  auto v = [](const double* p) { return mk(OC(p[0]), OC(p[1]), OC(p[2])); };
  V3<OC> f = self_force(v(tv), v(b), OC(sc[0]), OC(sc[1]), OC(sc[2]), OC(sc[3]), OC(sc[4]),
                        OC(sc[5]));
  (void)f;
  emit(counts);
}
}
