// CPU oracle for the DDD.jl force path: C entry points over the templated
// restatement in ddd_oracle.hpp.  TEST INFRASTRUCTURE ONLY -- loaded by tests/,
// the smoke check and the CPU-baseline leg of the benchmark, never by the
// product path.
//
// Array arguments use the reference's Julia memory layout (column-major,
// 1-based Int64 indices): coord 3 x nodeBuf, bVec 3 x linkBuf, segIdx
// ld x 3 (columns: link, node1, node2), extSeg one byte per segment,
// segForce 3 x 2 x ld, connectivity (1 + 2 maxConnect) x nodeBuf.
#include "ddd_oracle.hpp"

#include <atomic>
#include <cstring>
#include <vector>
#include <omp.h>

using namespace ddd_oracle;

namespace {

template <class T> SegConsts<T> consts_from(const double* K) {
  return SegConsts<T>{T(K[0]), T(K[1]), T(K[2]), T(K[3]), T(K[4]), T(K[5])};
}

struct SegView {
  int64_t numSeg;
  const int64_t* segIdx;
  int64_t ld;
  const double* coord;
  const double* bVec;
  const uint8_t* extSeg;
  // segment i (0-based row of segIdx)
  V3<double> b(int64_t i) const {
    const double* p = bVec + 3 * (segIdx[i] - 1);
    return mk(p[0], p[1], p[2]);
  }
  V3<double> n1(int64_t i) const {
    const double* p = coord + 3 * (segIdx[i + ld] - 1);
    return mk(p[0], p[1], p[2]);
  }
  V3<double> n2(int64_t i) const {
    const double* p = coord + 3 * (segIdx[i + 2 * ld] - 1);
    return mk(p[0], p[1], p[2]);
  }
  bool ext(int64_t i) const { return extSeg && extSeg[i]; }
};

template <class T> V3<T> cvt(const V3<double>& v) { return mk(T(v.x), T(v.y), T(v.z)); }

inline void add3(double* dst, const V3<double>& v) {
  dst[0] += v.x;
  dst[1] += v.y;
  dst[2] += v.z;
}
template <class T> inline void add3t(double* dst, const V3<T>& v) {
  dst[0] += double(v.x);
  dst[1] += double(v.y);
  dst[2] += double(v.z);
}

inline void atomic_add(double* p, double v) {
  // CAS loop: one atomic add per component, as Threads.atomic_add! does.
  auto* a = reinterpret_cast<std::atomic<double>*>(p);
  double old = a->load(std::memory_order_relaxed);
  while (!a->compare_exchange_weak(old, old + v, std::memory_order_relaxed)) {
  }
}

// calcSegSegForceSer!: DislocationForces.jl:778-829, in precision T,
// accumulated in T per segment then added into segForce.
template <class T>
void segseg_ser(const double* K, const SegView& s, double* segForce) {
  SegConsts<T> KK = consts_from<T>(K);
  const int64_t N = s.numSeg;
  std::vector<T> acc(6 * N);
  for (int64_t i = 0; i < 6 * N; ++i) acc[i] = T(segForce[i]);
  for (int64_t i = 0; i < N; ++i) {
    if (s.ext(i)) continue;
    V3<T> b1 = cvt<T>(s.b(i)), n11 = cvt<T>(s.n1(i)), n12 = cvt<T>(s.n2(i));
    for (int64_t j = i + 1; j < N; ++j) {
      if (s.ext(j)) continue;
      Forces4<T> f = seg_seg_force(KK, b1, n11, n12, cvt<T>(s.b(j)), cvt<T>(s.n1(j)),
                                   cvt<T>(s.n2(j)));
      T* ai = &acc[6 * i];
      T* aj = &acc[6 * j];
      ai[0] += f.f1.x; ai[1] += f.f1.y; ai[2] += f.f1.z;
      ai[3] += f.f2.x; ai[4] += f.f2.y; ai[5] += f.f2.z;
      aj[0] += f.f3.x; aj[1] += f.f3.y; aj[2] += f.f3.z;
      aj[3] += f.f4.x; aj[4] += f.f4.y; aj[5] += f.f4.z;
    }
  }
  for (int64_t i = 0; i < 6 * N; ++i) segForce[i] = double(acc[i]);
}

}  // namespace

namespace ddd_oracle {
int g_glibc_libm = 0;
}

extern "C" {

int orc_version() { return 1; }

// 0: fdlibm-method log/atan (default, shared with the kernels); 1: glibc.
void orc_set_libm(int glibc) { g_glibc_libm = glibc; }
double orc_libm_log(double x) { return ddd_libm::log(x); }
double orc_libm_atan(double x) { return ddd_libm::atan(x); }
// Vector form: y[i] = log(x[i]) (which 0), atan(x[i]) (1), sqrt(x[i]) (2),
// x[2i] / x[2i+1] (3), with the general IEEE operations.
void orc_libm_vec(int which, const double* x, int64_t n, double* y) {
  for (int64_t i = 0; i < n; ++i)
    y[i] = which == 0 ? ddd_libm::log(x[i]) : which == 1 ? ddd_libm::atan(x[i])
         : which == 2 ? std::sqrt(x[i]) : x[2 * i] / x[2 * i + 1];
}

// One pair, F1..F4 in out[12].  prec: 0 double, 1 long double, 2 float.
// Returns 1 when the near-parallel branch was taken.
int orc_segseg_pair(int prec, const double* K, const double* b1, const double* n11,
                    const double* n12, const double* b2, const double* n21, const double* n22,
                    double* out) {
  auto run = [&](auto tag) {
    using T = decltype(tag);
    SegConsts<T> KK = consts_from<T>(K);
    auto v = [](const double* p) { return mk(T(p[0]), T(p[1]), T(p[2])); };
    Forces4<T> f = seg_seg_force(KK, v(b1), v(n11), v(n12), v(b2), v(n21), v(n22));
    const V3<T>* fs[4] = {&f.f1, &f.f2, &f.f3, &f.f4};
    for (int k = 0; k < 4; ++k) {
      out[3 * k] = double(fs[k]->x);
      out[3 * k + 1] = double(fs[k]->y);
      out[3 * k + 2] = double(fs[k]->z);
    }
    return seg_seg_is_parallel(v(n11), v(n12), v(n21), v(n22)) ? 1 : 0;
  };
  if (prec == 1) return run((long double)0);
  if (prec == 2) return run(0.0f);
  return run(0.0);
}

// Whole-network serial seg-seg (calcSegSegForceSer!, :778-829); adds into
// segForce (3 x 2 x ld).  prec as above.
void orc_segseg_ser(int prec, const double* K, int64_t numSeg, const int64_t* segIdx, int64_t ld,
                    const double* coord, const double* bVec, const uint8_t* extSeg,
                    double* segForce) {
  SegView s{numSeg, segIdx, ld, coord, bVec, extSeg};
  if (prec == 1) segseg_ser<long double>(K, s, segForce);
  else if (prec == 2) segseg_ser<float>(K, s, segForce);
  else segseg_ser<double>(K, s, segForce);
}

// calcSegSegForcePar (:719-777) as the reference designs it: nthreads static
// chunks of the outer index over the upper triangle, 12 atomic adds per pair.
void orc_segseg_par_faithful(const double* K, int64_t numSeg, const int64_t* segIdx, int64_t ld,
                             const double* coord, const double* bVec, const uint8_t* extSeg,
                             double* segForce, int nthreads) {
  SegView s{numSeg, segIdx, ld, coord, bVec, extSeg};
  SegConsts<double> KK = consts_from<double>(K);
  std::vector<double> par(6 * numSeg, 0.0);
  double* P = par.data();
#pragma omp parallel num_threads(nthreads)
  {
    int tid = omp_get_thread_num();
    int nt = omp_get_num_threads();
    int64_t start = (int64_t(tid) * numSeg) / nt;
    int64_t stop = (int64_t(tid + 1) * numSeg) / nt;
    for (int64_t i = start; i < stop; ++i) {
      if (s.ext(i)) continue;
      V3<double> b1 = s.b(i), n11 = s.n1(i), n12 = s.n2(i);
      for (int64_t j = i + 1; j < numSeg; ++j) {
        if (s.ext(j)) continue;
        Forces4<double> f = seg_seg_force(KK, b1, n11, n12, s.b(j), s.n1(j), s.n2(j));
        double* pi = P + 6 * i;
        double* pj = P + 6 * j;
        atomic_add(pi + 0, f.f1.x); atomic_add(pi + 1, f.f1.y); atomic_add(pi + 2, f.f1.z);
        atomic_add(pi + 3, f.f2.x); atomic_add(pi + 4, f.f2.y); atomic_add(pi + 5, f.f2.z);
        atomic_add(pj + 0, f.f3.x); atomic_add(pj + 1, f.f3.y); atomic_add(pj + 2, f.f3.z);
        atomic_add(pj + 3, f.f4.x); atomic_add(pj + 4, f.f4.y); atomic_add(pj + 5, f.f4.z);
      }
    }
  }
  for (int64_t i = 0; i < 6 * numSeg; ++i) segForce[i] += P[i];
}

// calcSegSegForcePar's design restricted to a sample of outer rows: the rows
// are split into nthreads static chunks, each row i is paired with every
// j > i, and the 12 force components of each pair are added atomically.
// Used to time the reference's threaded path on a bounded share of a large
// network; returns the number of pairs evaluated.
// A tuned CPU design for the same sample of outer rows: dynamic scheduling
// and per-thread accumulators summed at the end, no atomics.  Timed beside
// the reference-faithful design as the second CPU baseline.
int64_t orc_segseg_tuned_rows(const double* K, int64_t numSeg, const int64_t* segIdx, int64_t ld,
                              const double* coord, const double* bVec, const uint8_t* extSeg,
                              const int64_t* rows, int64_t nRows, double* segForce,
                              int nthreads) {
  SegView s{numSeg, segIdx, ld, coord, bVec, extSeg};
  SegConsts<double> KK = consts_from<double>(K);
  int64_t pairs = 0;
  std::vector<std::vector<double>> priv((size_t)nthreads);
#pragma omp parallel num_threads(nthreads) reduction(+ : pairs)
  {
    std::vector<double>& acc = priv[(size_t)omp_get_thread_num()];
    acc.assign((size_t)(6 * numSeg), 0.0);
#pragma omp for schedule(dynamic, 1)
    for (int64_t r = 0; r < nRows; ++r) {
      int64_t i = rows[r] - 1;
      if (s.ext(i)) continue;
      V3<double> b1 = s.b(i), n11 = s.n1(i), n12 = s.n2(i);
      double fi[6] = {0, 0, 0, 0, 0, 0};
      for (int64_t j = i + 1; j < numSeg; ++j) {
        if (s.ext(j)) continue;
        Forces4<double> f = seg_seg_force(KK, b1, n11, n12, s.b(j), s.n1(j), s.n2(j));
        fi[0] += f.f1.x; fi[1] += f.f1.y; fi[2] += f.f1.z;
        fi[3] += f.f2.x; fi[4] += f.f2.y; fi[5] += f.f2.z;
        double* pj = acc.data() + 6 * j;
        pj[0] += f.f3.x; pj[1] += f.f3.y; pj[2] += f.f3.z;
        pj[3] += f.f4.x; pj[4] += f.f4.y; pj[5] += f.f4.z;
        ++pairs;
      }
      for (int q = 0; q < 6; ++q) acc[(size_t)(6 * i + q)] += fi[q];
    }
  }
  for (const auto& acc : priv)
    for (size_t e = 0; e < acc.size(); ++e) segForce[e] += acc[e];
  return pairs;
}

int64_t orc_segseg_par_rows(const double* K, int64_t numSeg, const int64_t* segIdx, int64_t ld,
                            const double* coord, const double* bVec, const uint8_t* extSeg,
                            const int64_t* rows, int64_t nRows, double* segForce, int nthreads) {
  SegView s{numSeg, segIdx, ld, coord, bVec, extSeg};
  SegConsts<double> KK = consts_from<double>(K);
  int64_t pairs = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : pairs)
  {
    int tid = omp_get_thread_num();
    int nt = omp_get_num_threads();
    int64_t start = (int64_t(tid) * nRows) / nt;
    int64_t stop = (int64_t(tid + 1) * nRows) / nt;
    for (int64_t r = start; r < stop; ++r) {
      int64_t i = rows[r] - 1;
      if (s.ext(i)) continue;
      V3<double> b1 = s.b(i), n11 = s.n1(i), n12 = s.n2(i);
      for (int64_t j = i + 1; j < numSeg; ++j) {
        if (s.ext(j)) continue;
        Forces4<double> f = seg_seg_force(KK, b1, n11, n12, s.b(j), s.n1(j), s.n2(j));
        double* pi = segForce + 6 * i;
        double* pj = segForce + 6 * j;
        atomic_add(pi + 0, f.f1.x); atomic_add(pi + 1, f.f1.y); atomic_add(pi + 2, f.f1.z);
        atomic_add(pi + 3, f.f2.x); atomic_add(pi + 4, f.f2.y); atomic_add(pi + 5, f.f2.z);
        atomic_add(pj + 0, f.f3.x); atomic_add(pj + 1, f.f3.y); atomic_add(pj + 2, f.f3.z);
        atomic_add(pj + 3, f.f4.x); atomic_add(pj + 4, f.f4.y); atomic_add(pj + 5, f.f4.z);
        ++pairs;
      }
    }
  }
  return pairs;
}

// The forces of the all-pairs triangle on sampled segments: for row
// i = rows[k] - 1, the sum over j != i in index order of the pair force on
// segment i in the triangle's orientation (calcSegSegForceSer!,
// DislocationForces.jl:794-825: F1, F2 of pair (i, j) for j > i, F3, F4 of
// pair (j, i) for j < i).  out: 6 x nRows.
void orc_segseg_rows_triangle(const double* K, int64_t numSeg, const int64_t* segIdx, int64_t ld,
                              const double* coord, const double* bVec, const uint8_t* extSeg,
                              const int64_t* rows, int64_t nRows, double* out, int nthreads) {
  SegView s{numSeg, segIdx, ld, coord, bVec, extSeg};
  SegConsts<double> KK = consts_from<double>(K);
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
  for (int64_t r = 0; r < nRows; ++r) {
    const int64_t i = rows[r] - 1;
    double fi[6] = {0, 0, 0, 0, 0, 0};
    if (!s.ext(i)) {
      for (int64_t j = 0; j < numSeg; ++j) {
        if (j == i || s.ext(j)) continue;
        V3<double> a, b;
        if (j > i) {
          Forces4<double> f = seg_seg_force(KK, s.b(i), s.n1(i), s.n2(i), s.b(j), s.n1(j), s.n2(j));
          a = f.f1;
          b = f.f2;
        } else {
          Forces4<double> f = seg_seg_force(KK, s.b(j), s.n1(j), s.n2(j), s.b(i), s.n1(i), s.n2(i));
          a = f.f3;
          b = f.f4;
        }
        fi[0] += a.x; fi[1] += a.y; fi[2] += a.z;
        fi[3] += b.x; fi[4] += b.y; fi[5] += b.z;
      }
    }
    for (int q = 0; q < 6; ++q) out[6 * r + q] = fi[q];
  }
}

// Rows-only forces split by distance: for each row i = rows[k]-1 against
// every j != i (row segment first), the F1/F2 contributions are summed into
// bin b where edges[b] <= |mid_i - mid_j| < edges[b+1] (edges ascending,
// nb+1 values).  out: 6 x nb x nRows.  Analysis of how a node's force is
// distributed over source distance.
void orc_segseg_rows_binned(const double* K, int64_t numSeg, const int64_t* segIdx, int64_t ld,
                            const double* coord, const double* bVec, const uint8_t* extSeg,
                            const int64_t* rows, int64_t nRows, const double* edges, int nb,
                            double* out, int nthreads) {
  SegView s{numSeg, segIdx, ld, coord, bVec, extSeg};
  SegConsts<double> KK = consts_from<double>(K);
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
  for (int64_t k = 0; k < nRows; ++k) {
    int64_t i = rows[k] - 1;
    if (s.ext(i)) continue;
    V3<double> b1 = s.b(i), n11 = s.n1(i), n12 = s.n2(i);
    V3<double> m1 = (n11 + n12) * 0.5;
    double* o = out + 6 * nb * k;
    for (int64_t j = 0; j < numSeg; ++j) {
      if (j == i || s.ext(j)) continue;
      V3<double> n21 = s.n1(j), n22 = s.n2(j);
      double dist = norm((n21 + n22) * 0.5 - m1);
      int b = 0;
      while (b < nb && !(dist < edges[b + 1])) ++b;
      if (b >= nb || dist < edges[0]) continue;
      Forces4<double> f = seg_seg_force(KK, b1, n11, n12, s.b(j), n21, n22);
      double* ob = o + 6 * b;
      ob[0] += f.f1.x; ob[1] += f.f1.y; ob[2] += f.f1.z;
      ob[3] += f.f2.x; ob[4] += f.f2.y; ob[5] += f.f2.z;
    }
  }
}

// Rows-only forces split by source group: group[j] in [0, ng) names the
// group of source segment j (negative: skipped).  out: 6 x ng x nRows.
void orc_segseg_rows_grouped(const double* K, int64_t numSeg, const int64_t* segIdx, int64_t ld,
                             const double* coord, const double* bVec, const uint8_t* extSeg,
                             const int64_t* rows, int64_t nRows, const int64_t* group, int64_t ng,
                             double* out, int nthreads) {
  SegView s{numSeg, segIdx, ld, coord, bVec, extSeg};
  SegConsts<double> KK = consts_from<double>(K);
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
  for (int64_t k = 0; k < nRows; ++k) {
    int64_t i = rows[k] - 1;
    if (s.ext(i)) continue;
    V3<double> b1 = s.b(i), n11 = s.n1(i), n12 = s.n2(i);
    double* o = out + 6 * ng * k;
    for (int64_t j = 0; j < numSeg; ++j) {
      if (j == i || s.ext(j) || group[j] < 0) continue;
      Forces4<double> f = seg_seg_force(KK, b1, n11, n12, s.b(j), s.n1(j), s.n2(j));
      double* ob = o + 6 * group[j];
      ob[0] += f.f1.x; ob[1] += f.f1.y; ob[2] += f.f1.z;
      ob[3] += f.f2.x; ob[4] += f.f2.y; ob[5] += f.f2.z;
    }
  }
}

// Mixed-precision whole-network seg-seg, the rule of the mixed path: pairs
// with midpoint distance > farDist and |t1 x t2| > sinMin use the far-pair
// quadrature in precision prec (2 float, 0 double) with nq points; all other
// pairs the FP64 analytic force.  Accumulation in double.  Returns the number
// of pairs on the FP64 path.
int64_t orc_segseg_mixed(int prec, int nq, double farDist, double sinMin, const double* K,
                         const double* sc, int64_t numSeg, const int64_t* segIdx, int64_t ld,
                         const double* coord, const double* bVec, const uint8_t* extSeg,
                         double* segForce, int nthreads) {
  SegView s{numSeg, segIdx, ld, coord, bVec, extSeg};
  SegConsts<double> KK = consts_from<double>(K);
  int64_t exact = 0;
  std::vector<double> acc(6 * numSeg, 0.0);
  double* A = acc.data();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 16) reduction(+ : exact)
  for (int64_t i = 0; i < numSeg; ++i) {
    if (s.ext(i)) continue;
    V3<double> b1 = s.b(i), n11 = s.n1(i), n12 = s.n2(i);
    V3<double> t1 = (n12 - n11) / norm(n12 - n11), m1 = (n11 + n12) * 0.5;
    for (int64_t j = i + 1; j < numSeg; ++j) {
      if (s.ext(j)) continue;
      V3<double> b2 = s.b(j), n21 = s.n1(j), n22 = s.n2(j);
      V3<double> t2 = (n22 - n21) / norm(n22 - n21);
      double dist = norm((n21 + n22) * 0.5 - m1);
      bool far = dist > farDist && norm(cross(t1, t2)) > sinMin;
      V3<double> F1, F2, F3, F4;
      if (far) {
        if (prec == 2) {
          far_side_quad<float>(b1, n11, n12, b2, n21, n22, nq, float(sc[0]), float(sc[1]), float(sc[2]), float(sc[3]), float(sc[4]), F1, F2);
          far_side_quad<float>(b2, n21, n22, b1, n11, n12, nq, float(sc[0]), float(sc[1]), float(sc[2]), float(sc[3]), float(sc[4]), F3, F4);
        } else {
          far_side_quad<double>(b1, n11, n12, b2, n21, n22, nq, sc[0], sc[1], sc[2], sc[3], sc[4], F1, F2);
          far_side_quad<double>(b2, n21, n22, b1, n11, n12, nq, sc[0], sc[1], sc[2], sc[3], sc[4], F3, F4);
        }
      } else {
        Forces4<double> f = seg_seg_force(KK, b1, n11, n12, b2, n21, n22);
        F1 = f.f1; F2 = f.f2; F3 = f.f3; F4 = f.f4;
        ++exact;
      }
      double* ai = A + 6 * i;
      double* aj = A + 6 * j;
      atomic_add(ai + 0, F1.x); atomic_add(ai + 1, F1.y); atomic_add(ai + 2, F1.z);
      atomic_add(ai + 3, F2.x); atomic_add(ai + 4, F2.y); atomic_add(ai + 5, F2.z);
      atomic_add(aj + 0, F3.x); atomic_add(aj + 1, F3.y); atomic_add(aj + 2, F3.z);
      atomic_add(aj + 3, F4.x); atomic_add(aj + 4, F4.y); atomic_add(aj + 5, F4.z);
    }
  }
  for (int64_t i = 0; i < 6 * numSeg; ++i) segForce[i] += A[i];
  return exact;
}

// Near-field rows of the hybrid split: for every live segment i, the F1/F2
// of every pair (i, j), j != i, whose cells (cellOf, x + G (y + G z)) are
// within Chebyshev distance rn, segment i first, in FP64; added to out.
void orc_segseg_near(const double* K, int64_t numSeg, const int64_t* segIdx, int64_t ld,
                     const double* coord, const double* bVec, const uint8_t* extSeg,
                     const int64_t* cellOf, int G, int rn, double* out, int nthreads) {
  SegView s{numSeg, segIdx, ld, coord, bVec, extSeg};
  SegConsts<double> KK = consts_from<double>(K);
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 16)
  for (int64_t i = 0; i < numSeg; ++i) {
    if (s.ext(i) || cellOf[i] < 0) continue;
    const int64_t ci = cellOf[i];
    const int ix = int(ci % G), iy = int((ci / G) % G), iz = int(ci / (int64_t(G) * G));
    V3<double> b1 = s.b(i), n11 = s.n1(i), n12 = s.n2(i);
    double* o = out + 6 * i;
    for (int64_t j = 0; j < numSeg; ++j) {
      if (j == i || s.ext(j) || cellOf[j] < 0) continue;
      const int64_t cj = cellOf[j];
      const int jx = int(cj % G), jy = int((cj / G) % G), jz = int(cj / (int64_t(G) * G));
      if (std::abs(jx - ix) > rn || std::abs(jy - iy) > rn || std::abs(jz - iz) > rn) continue;
      Forces4<double> f = seg_seg_force(KK, b1, n11, n12, s.b(j), s.n1(j), s.n2(j));
      add3(o, f.f1);
      add3(o + 3, f.f2);
    }
  }
}

// The same near pairs, each evaluated once with the segment of lower rank as
// the first segment and its forces added to both segments (the pair
// orientation of a kernel that walks segments in rank order).  Serial.
void orc_segseg_near_once(const double* K, int64_t numSeg, const int64_t* segIdx, int64_t ld,
                          const double* coord, const double* bVec, const uint8_t* extSeg,
                          const int64_t* cellOf, int G, int rn, const int64_t* rank, double* out) {
  SegView s{numSeg, segIdx, ld, coord, bVec, extSeg};
  SegConsts<double> KK = consts_from<double>(K);
  for (int64_t i = 0; i < numSeg; ++i) {
    if (s.ext(i) || cellOf[i] < 0) continue;
    const int64_t ci = cellOf[i];
    const int ix = int(ci % G), iy = int((ci / G) % G), iz = int(ci / (int64_t(G) * G));
    for (int64_t j = 0; j < numSeg; ++j) {
      if (rank[j] <= rank[i] || s.ext(j) || cellOf[j] < 0) continue;
      const int64_t cj = cellOf[j];
      const int jx = int(cj % G), jy = int((cj / G) % G), jz = int(cj / (int64_t(G) * G));
      if (std::abs(jx - ix) > rn || std::abs(jy - iy) > rn || std::abs(jz - iz) > rn) continue;
      Forces4<double> f = seg_seg_force(KK, s.b(i), s.n1(i), s.n2(i), s.b(j), s.n1(j), s.n2(j));
      add3(out + 6 * i, f.f1);
      add3(out + 6 * i + 3, f.f2);
      add3(out + 6 * j, f.f3);
      add3(out + 6 * j + 3, f.f4);
    }
  }
}

// Rows-only seg-seg over a subset (calcSegSegForceIdx!, :831-877).  For each
// k, row i = idx[k]-1 is computed against every j != i with segment i as the
// first segment; F1/F2 are added to out column (byPosition ? k : i).
void orc_segseg_idx(const double* K, int64_t numSeg, const int64_t* segIdx, int64_t ld,
                    const double* coord, const double* bVec, const uint8_t* extSeg,
                    const int64_t* idx, int64_t nIdx, double* out, int byPosition) {
  SegView s{numSeg, segIdx, ld, coord, bVec, extSeg};
  SegConsts<double> KK = consts_from<double>(K);
  for (int64_t k = 0; k < nIdx; ++k) {
    int64_t i = idx[k] - 1;
    if (s.ext(i)) continue;
    V3<double> b1 = s.b(i), n11 = s.n1(i), n12 = s.n2(i);
    double* o = out + 6 * (byPosition ? k : i);
    for (int64_t j = 0; j < numSeg; ++j) {
      if (j == i || s.ext(j)) continue;
      Forces4<double> f = seg_seg_force(KK, b1, n11, n12, s.b(j), s.n1(j), s.n2(j));
      add3(o, f.f1);
      add3(o + 3, f.f2);
    }
  }
}

// Self force (calcSelfForce!, :354-388 and _calcSelfForce :414-455).
// sc = {a, aSq, mu4pi, nuOmNuInv, Ec, omNuInv}.  idx NULL means all
// segments.  Node 1 receives -f, node 2 +f, at out column
// (byPosition ? k : i).
void orc_self_force(const double* sc, int64_t numSeg, const int64_t* segIdx, int64_t ld,
                    const double* coord, const double* bVec, const uint8_t* extSeg,
                    const int64_t* idx, int64_t nIdx, double* out, int byPosition) {
  SegView s{numSeg, segIdx, ld, coord, bVec, extSeg};
  int64_t n = idx ? nIdx : numSeg;
  for (int64_t k = 0; k < n; ++k) {
    int64_t i = idx ? idx[k] - 1 : k;
    if (s.ext(i)) continue;
    V3<double> f = self_force(s.n2(i) - s.n1(i), s.b(i), sc[0], sc[1], sc[2], sc[3], sc[4], sc[5]);
    double* o = out + 6 * (byPosition ? k : i);
    o[0] -= f.x; o[1] -= f.y; o[2] -= f.z;
    o[3] += f.x; o[4] += f.y; o[5] += f.z;
  }
}

// Stress from the FE correction field at x0 (calc_sigmaHat,
// DislocationForces.jl:92-192; shapeFunctionDeriv FEMShapeFunction.jl:83-122).
// C is 6x6 column-major, mcoord 3 x numNode, conn 8 x numElem (1-based),
// uHat dense (3 numNode).  out = 3x3 column-major.
void orc_sigma_hat(const double* C, const double* mcoord, const int64_t* conn, int64_t mx,
                   int64_t my, int64_t mz, double w, double h, double d, const double* uHat,
                   const double* x0, double* out) {
  double wInv = 1 / w, hInv = 1 / h, dInv = 1 / d;
  double x = x0[0], y = x0[1], z = x0[2];
  auto clampi = [](double v, int64_t lo, int64_t hi) {
    double c = std::ceil(v);
    if (c < lo) c = double(lo);
    if (c > hi) c = double(hi);
    return int64_t(c);
  };
  int64_t i = clampi(x * wInv, 1, mx), j = clampi(y * hInv, 1, my), k = clampi(z * dInv, 1, mz);
  int64_t e = i + (j - 1) * mx + (k - 1) * mx * my;  // 1-based element
  const int64_t* ce = conn + 8 * (e - 1);
  int64_t nn1 = ce[0], nn7 = ce[6];
  double xc = 0.5 * (mcoord[3 * (nn1 - 1)] + mcoord[3 * (nn7 - 1)]);
  double yc = 0.5 * (mcoord[3 * (nn1 - 1) + 1] + mcoord[3 * (nn7 - 1) + 1]);
  double zc = 0.5 * (mcoord[3 * (nn1 - 1) + 2] + mcoord[3 * (nn7 - 1) + 2]);
  double ds1dx = 2 * wInv, ds2dy = 2 * hInv, ds3dz = 2 * dInv;
  double s1 = (x - xc) * ds1dx, s2 = (y - yc) * ds2dy, s3 = (z - zc) * ds3dz;
  double omx = 1 - s1, omy = 1 - s2, omz = 1 - s3, opx = 1 + s1, opy = 1 + s2, opz = 1 + s3;
  // dNdS[r][n], r = derivative direction, n = local node.
  double dN[3][8] = {
      {-omy * omz, omy * omz, opy * omz, -opy * omz, -omy * opz, omy * opz, opy * opz, -opy * opz},
      {-omx * omz, -opx * omz, opx * omz, omx * omz, -omx * opz, -opx * opz, opx * opz, omx * opz},
      {-omx * omy, -opx * omy, -opx * opy, -omx * opy, omx * omy, opx * omy, opx * opy, omx * opy}};
  for (int r = 0; r < 3; ++r)
    for (int n = 0; n < 8; ++n) dN[r][n] = dN[r][n] / 8;
  double B[6][24];
  double U[24];
  std::memset(B, 0, sizeof(B));
  for (int n = 1; n <= 8; ++n) {
    int idx1 = 3 * n, idx2 = 3 * (n - 1);
    int64_t idx3 = 3 * ce[n - 1];
    B[0][idx2 + 0] = dN[0][n - 1] * ds1dx;
    B[1][idx2 + 1] = dN[1][n - 1] * ds2dy;
    B[2][idx2 + 2] = dN[2][n - 1] * ds3dz;
    B[3][idx2 + 0] = B[1][idx2 + 1];
    B[3][idx2 + 1] = B[0][idx2 + 0];
    B[4][idx2 + 0] = B[2][idx1 - 1];
    B[4][idx1 - 1] = B[0][idx2 + 0];
    B[5][idx2 + 1] = B[2][idx1 - 1];
    B[5][idx1 - 1] = B[1][idx2 + 1];
    U[idx1 - 3] = uHat[idx3 - 3];
    U[idx1 - 2] = uHat[idx3 - 2];
    U[idx1 - 1] = uHat[idx3 - 1];
  }
  double eps6[6], sv[6];
  for (int r = 0; r < 6; ++r) {
    double acc = 0;
    for (int c = 0; c < 24; ++c) acc += B[r][c] * U[c];
    eps6[r] = acc;
  }
  for (int r = 0; r < 6; ++r) {
    double acc = 0;
    for (int c = 0; c < 6; ++c) acc += C[r + 6 * c] * eps6[c];
    sv[r] = acc;
  }
  // Voigt (xx, yy, zz, xy, xz, yz) -> 3x3 column-major.
  double S[9] = {sv[0], sv[3], sv[4], sv[3], sv[1], sv[5], sv[4], sv[5], sv[2]};
  std::memcpy(out, S, sizeof(S));
}

// PK force (calcPKForce!, :244-300): pk = (sigmaHat(mid) b) x (n2 - n1);
// half to each node at out column (byPosition ? k : i).
void orc_pk_force(const double* C, const double* mcoord, const int64_t* conn, int64_t mx,
                  int64_t my, int64_t mz, double w, double h, double d, const double* uHat,
                  int64_t numSeg, const int64_t* segIdx, int64_t ld, const double* coord,
                  const double* bVec, const uint8_t* extSeg, const int64_t* idx, int64_t nIdx,
                  double* out, int byPosition, int halfToNodes) {
  SegView s{numSeg, segIdx, ld, coord, bVec, extSeg};
  int64_t n = idx ? nIdx : numSeg;
  for (int64_t k = 0; k < n; ++k) {
    int64_t i = idx ? idx[k] - 1 : k;
    if (s.ext(i)) continue;
    V3<double> a = s.n1(i), bb = s.n2(i), b = s.b(i);
    V3<double> t = bb - a;
    V3<double> mid = (bb + a) / 2.0;
    double x0[3] = {mid.x, mid.y, mid.z}, S[9];
    orc_sigma_hat(C, mcoord, conn, mx, my, mz, w, h, d, uHat, x0, S);
    V3<double> sb = mk(S[0] * b.x + S[3] * b.y + S[6] * b.z, S[1] * b.x + S[4] * b.y + S[7] * b.z,
                       S[2] * b.x + S[5] * b.y + S[8] * b.z);
    V3<double> pk = cross(sb, t);
    if (halfToNodes) {
      double* o = out + 6 * (byPosition ? k : i);
      o[0] += pk.x / 2; o[1] += pk.y / 2; o[2] += pk.z / 2;
      o[3] += pk.x / 2; o[4] += pk.y / 2; o[5] += pk.z / 2;
    } else {
      double* o = out + 3 * (byPosition ? k : i);
      o[0] += pk.x; o[1] += pk.y; o[2] += pk.z;
    }
  }
}

// Field-point stress (_calc_sigmaTilde!, DislocationFEM.jl:86-303):
// sc = {a2, mu4pi, mu4pinu, nuMu4pinu, a2mu8pi}; x0 3 x P; sigma 6 x P
// accumulated.  Does not skip external segments, as the reference does not.
void orc_sigma_tilde(const double* sc, int64_t numSeg, const int64_t* segIdx, int64_t ld,
                     const double* coord, const double* bVec, const int64_t* idx, int64_t nIdx,
                     const double* x0, int64_t P, double* sigma, int nthreads) {
  SegView s{numSeg, segIdx, ld, coord, bVec, nullptr};
  int64_t n = idx ? nIdx : numSeg;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t p = 0; p < P; ++p) {
    V3<double> x = mk(x0[3 * p], x0[3 * p + 1], x0[3 * p + 2]);
    double* sg = sigma + 6 * p;
    for (int64_t k = 0; k < n; ++k) {
      int64_t i = idx ? idx[k] - 1 : k;
      sigma_tilde_point(s.b(i), s.n1(i), s.n2(i), x, sc[0], sc[1], sc[2], sc[3], sc[4], sg);
    }
  }
}

// Node-force gather (dlnMobility!, DislocationMobility.jl:220-255):
// nodeForce[:, n] = sum over connections of segForce[:, colLink, link],
// skipping zero-length links and contributions below sigmaPN.
void orc_node_gather(const int64_t* connectivity, int64_t ldConn, int64_t numNode,
                     const int64_t* links, const double* coord, const double* segForce,
                     double sigmaPN, double* nodeForce) {
  const double eps = std::numeric_limits<double>::epsilon();
  for (int64_t n = 0; n < numNode; ++n) {
    const int64_t* cn = connectivity + ldConn * n;
    double f[3] = {0, 0, 0};
    for (int64_t c = 1; c <= cn[0]; ++c) {
      int64_t link = cn[2 * c - 1];
      int64_t col = cn[2 * c];
      int64_t node2 = links[2 * (link - 1) + (3 - col) - 1];
      double tx = coord[3 * (node2 - 1)] - coord[3 * n];
      double ty = coord[3 * (node2 - 1) + 1] - coord[3 * n + 1];
      double tz = coord[3 * (node2 - 1) + 2] - coord[3 * n + 2];
      double nT = std::sqrt(tx * tx + ty * ty + tz * tz);
      if (nT < eps) continue;
      double tN = 1 / nT;
      const double* sf = segForce + 6 * (link - 1) + 3 * (col - 1);
      double fx = sf[0], fy = sf[1], fz = sf[2];
      if (std::sqrt(fx * fx + fy * fy + fz * fz) * tN < sigmaPN) fx = fy = fz = 0;
      f[0] += fx; f[1] += fy; f[2] += fz;
    }
    nodeForce[3 * n] = f[0];
    nodeForce[3 * n + 1] = f[1];
    nodeForce[3 * n + 2] = f[2];
  }
}

}  // extern "C"

extern "C" {

// BCC mobility of the nodes idx[0..nIdx) (all numNode nodes when idx is
// NULL): dlnMobility / dlnMobility! for mobBCC, DislocationMobility.jl:16-171
// and :186-341.  drag = (edge, screw, climb, line), the reference's
// dragCoeffs.  Output column k belongs to the k-th node of the list;
// dragOut (9 doubles per node, column-major, may be NULL) receives the drag
// tensor the velocity solves with.
//
// Per connection (:56-146): skip links shorter than eps (:67-71); the force
// column of the connection, zeroed below sigmaPN (:74-80); then the drag of
// the link: an edge-type law climb I + (line - climb) t t when b is not a
// <111>-family screw vector (:84-114), else the screw law screw I +
// (line - screw) t t plus, unless the link is pure screw, the glide and
// climb corrections along q and p (:116-145, Arsenlis et al. 2007 eq. 112).
// Velocity (:148): the 3x3 system drag v = f solved by Cramer's rule, which
// is what the reference's StaticArrays `\` does for a 3x3 SMatrix
// (StaticArrays.jl, no version pinned): here as the adjugate built from
// cross products of the columns over the determinant as a triple product.
// The rounding of that solve therefore differs from Julia's by a few ulp
// times the condition number of the drag (up to ~1e10 with climb = 1e10);
// the reference's own test allows rtol 1e-5 (MobilityTest.jl:92).
void orc_mobility_bcc(const double* drag, double sigmaPN, const int64_t* connectivity,
                      int64_t ldConn, int64_t numNode, const int64_t* links, const double* coord,
                      const double* bVec, const double* segForce, const int64_t* idx, int64_t nIdx,
                      double* nodeForce, double* nodeVel, double* dragOut) {
  const double eps = std::numeric_limits<double>::epsilon();
  const double sqrtEps = std::sqrt(eps);
  const double edgeD = drag[0], screwD = drag[1], climbD = drag[2], lineD = drag[3];
  const int64_t n = idx ? nIdx : numNode;
  for (int64_t k = 0; k < n; ++k) {
    const int64_t node = idx ? idx[k] - 1 : k;
    const int64_t* cn = connectivity + ldConn * node;
    double D[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double f[3] = {0, 0, 0};
    const V3<double> x0 = mk(coord[3 * node], coord[3 * node + 1], coord[3 * node + 2]);
    for (int64_t c = 1; c <= cn[0]; ++c) {
      const int64_t link = cn[2 * c - 1], col = cn[2 * c];
      const int64_t node2 = links[2 * (link - 1) + (3 - col) - 1];
      V3<double> t = mk(coord[3 * (node2 - 1)], coord[3 * (node2 - 1) + 1],
                        coord[3 * (node2 - 1) + 2]) - x0;
      const double nT = norm(t);
      if (nT < eps) continue;
      const double tN = 1 / nT;
      t = t * tN;
      const double* sf = segForce + 6 * (link - 1) + 3 * (col - 1);
      V3<double> fc = mk(sf[0], sf[1], sf[2]);
      if (norm(fc) * tN < sigmaPN) fc = mk(0.0, 0.0, 0.0);
      f[0] += fc.x; f[1] += fc.y; f[2] += fc.z;

      V3<double> b = mk(bVec[3 * (link - 1)], bVec[3 * (link - 1) + 1], bVec[3 * (link - 1) + 2]);
      const double nB = norm(b);
      // <111>-family test on |b| / min|b_i| (zero components give Inf/NaN
      // ratios that match no family, as in the reference).
      const double ab[3] = {std::fabs(b.x), std::fabs(b.y), std::fabs(b.z)};
      const double mn = std::min(ab[0], std::min(ab[1], ab[2]));
      int n1 = 0, n2 = 0, n3 = 0;
      for (int q = 0; q < 3; ++q) {
        const double r = ab[q] / mn;
        n1 += std::fabs(1 - r) < sqrtEps;
        n2 += std::fabs(2 - r) < sqrtEps;
        n3 += std::fabs(3 - r) < sqrtEps;
      }
      const bool unitLike = 1 - nB < sqrtEps;
      b = b / nB;
      const double w = nB * nT / 2;
      const double tv[3] = {t.x, t.y, t.z};
      const bool screwFamily = unitLike && ((n1 == 3) || (n1 == 2 && n2 == 1) ||
                                            (n1 == 1 && n2 == 1 && n3 == 1));
      const double base = screwFamily ? screwD : climbD;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          D[i][j] += w * ((i == j ? base : 0.0) + (lineD - base) * tv[i] * tv[j]);
      if (!screwFamily) continue;
      const double cosSq = dot(t, b) * dot(t, b);
      const double sinSq = 1 - cosSq;
      if (sinSq < eps) continue;
      const V3<double> p = cross(b, t) / std::sqrt(sinSq);
      const V3<double> q = cross(p, t);
      const double glide = 1 / std::sqrt(sinSq / (edgeD * edgeD) + cosSq / (screwD * screwD));
      const double climbS = std::sqrt(sinSq * (climbD * climbD) + cosSq * (screwD * screwD));
      const double pv[3] = {p.x, p.y, p.z}, qv[3] = {q.x, q.y, q.z};
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          D[i][j] += w * ((glide - screwD) * qv[i] * qv[j] + (climbS - screwD) * pv[i] * pv[j]);
    }
    // Cramer's rule: inverse rows are the cross products of the columns.
    const V3<double> c0 = mk(D[0][0], D[1][0], D[2][0]);
    const V3<double> c1 = mk(D[0][1], D[1][1], D[2][1]);
    const V3<double> c2 = mk(D[0][2], D[1][2], D[2][2]);
    const V3<double> r0 = cross(c1, c2), r1 = cross(c2, c0), r2 = cross(c0, c1);
    const double det = dot(c0, r0);
    const V3<double> fv = mk(f[0], f[1], f[2]);
    nodeForce[3 * k] = f[0];
    nodeForce[3 * k + 1] = f[1];
    nodeForce[3 * k + 2] = f[2];
    nodeVel[3 * k] = dot(r0, fv) / det;
    nodeVel[3 * k + 1] = dot(r1, fv) / det;
    nodeVel[3 * k + 2] = dot(r2, fv) / det;
    if (dragOut)
      for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) dragOut[9 * k + 3 * j + i] = D[i][j];
  }
}

}  // extern "C"

// ---- calc_uTilde: dislocation displacements at FE nodes -------------------
// DislocationFEM.jl:355-600 (_calc_uTilde!), the triangle displacement of
// :602-707 (Barnett triangles, Bromage & Tarleton 2018), safeNorm
// (Misc.jl:289-299), linePlaneIntersect (Misc.jl:223-235) and
// findIntersectVolume (DislocationTopologySpecial.jl:7-38).
// Internal segments are pinned by the reference's golden displacements.  The
// closure of segments with an external node (their ends projected onto the
// mesh surface, DislocationFEM.jl:437-528), for which the reference's tests
// hold no vectors, is pinned by geometry: the surface points in closed form
// (and one hand-worked trace of the face-plane walk), the polygon's solid
// angle by another triangulation and formula, up to whole Burgers vectors.
namespace {

struct UT {
  double omNuInv8pi, om2NuOmNuInv8pi;
};

inline double sq(double x) { return x * x; }

// safeNorm: (|x|, x/|x|), or (eps, 0) below eps.
inline V3<double> safe_norm(const V3<double>& x, double& n) {
  const double eps = std::numeric_limits<double>::epsilon();
  n = norm(x);
  if (n > eps) return mk(x.x / n, x.y / n, x.z / n);
  n = eps;
  return mk(0.0, 0.0, 0.0);
}

inline double clamp1(double x) { return x < -1 ? -1 : (x > 1 ? 1 : x); }
inline double sgn(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0); }

// Solid angle of triangle ABC seen from the point, by l'Huilier's theorem
// on the unit vectors towards the corners, signed by RA . (CA x AB).
inline double solid_angle(const V3<double>& RA, const V3<double>& RB, const V3<double>& RC,
                          const V3<double>& AB, const V3<double>& CA) {
  const double ab = clamp1(dot(RA, RB)), bc = clamp1(dot(RB, RC)), ca = clamp1(dot(RC, RA));
  const double ta = std::acos(ab), tb = std::acos(bc), tc = std::acos(ca);
  const double t = (ta + tb + tc) / 2;
  double f = std::tan(t / 2) * std::tan((t - ta) / 2) * std::tan((t - tb) / 2) *
             std::tan((t - tc) / 2);
  f = std::max(f, 0.0);
  const double w = 4 * std::atan(std::sqrt(f));
  return -sgn(dot(RA, cross(CA, AB))) * w;
}

// calcDisplacementDislocationTriangle! with the material terms (:602-670).
void tri_full(const UT& k, const V3<double>& A, const V3<double>& B, const V3<double>& C,
              const V3<double>& b, const double* pts, int64_t P, double* u) {
  double dummy;
  const V3<double> AB = safe_norm(B - A, dummy), BC = safe_norm(C - B, dummy),
                   CA = safe_norm(A - C, dummy);
  const V3<double> bAB = cross(b, AB), bBC = cross(b, BC), bCA = cross(b, CA);
  const double fourPi = 4 * M_PI;
  for (int64_t i = 0; i < P; ++i) {
    const V3<double> p = mk(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
    double ra, rb, rc;
    const V3<double> RA = safe_norm(A - p, ra), RB = safe_norm(B - p, rb), RC = safe_norm(C - p, rc);
    const V3<double> fAB = bAB * std::log((rb / ra) * (1 + dot(RB, AB)) / (1 + dot(RA, AB)));
    const V3<double> fBC = bBC * std::log((rc / rb) * (1 + dot(RC, BC)) / (1 + dot(RB, BC)));
    const V3<double> fCA = bCA * std::log((ra / rc) * (1 + dot(RA, CA)) / (1 + dot(RC, CA)));
    const double ab = clamp1(dot(RA, RB)), bc = clamp1(dot(RB, RC)), ca = clamp1(dot(RC, RA));
    const V3<double> gAB = (dot(b, cross(RA, RB)) * (RA + RB)) / (1 + ab);
    const V3<double> gBC = (dot(b, cross(RB, RC)) * (RB + RC)) / (1 + bc);
    const V3<double> gCA = (dot(b, cross(RC, RA)) * (RC + RA)) / (1 + ca);
    const double w = solid_angle(RA, RB, RC, AB, CA);
    const V3<double> ui = (((-b) * w) / fourPi - k.om2NuOmNuInv8pi * ((fAB + fBC) + fCA)) +
                          k.omNuInv8pi * ((gAB + gBC) + gCA);
    u[3 * i] += ui.x;
    u[3 * i + 1] += ui.y;
    u[3 * i + 2] += ui.z;
  }
}

// The solid-angle term alone (:671-707), for the closures of external
// segments.
void tri_solid(const V3<double>& A, const V3<double>& B, const V3<double>& C, const V3<double>& b,
               const double* pts, int64_t P, double* u) {
  double dummy;
  const V3<double> AB = safe_norm(B - A, dummy), CA = safe_norm(A - C, dummy);
  const double fourPi = 4 * M_PI;
  for (int64_t i = 0; i < P; ++i) {
    const V3<double> p = mk(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
    double r;
    const V3<double> RA = safe_norm(A - p, r), RB = safe_norm(B - p, r), RC = safe_norm(C - p, r);
    const double w = solid_angle(RA, RB, RC, AB, CA);
    const V3<double> ui = ((-b) * w) / fourPi;
    u[3 * i] += ui.x;
    u[3 * i + 1] += ui.y;
    u[3 * i + 2] += ui.z;
  }
}

struct Cuboid {
  const double* faceNorm;   // 3 x 6
  const double* faceMidPt;  // 3 x 6
  double lo[3], hi[3];      // bounding box of the vertices
  V3<double> n(int i) const { return mk(faceNorm[3 * i], faceNorm[3 * i + 1], faceNorm[3 * i + 2]); }
  V3<double> m(int i) const {
    return mk(faceMidPt[3 * i], faceMidPt[3 * i + 1], faceMidPt[3 * i + 2]);
  }
  // membership of the convex hull of the vertices (a box here), with an
  // absolute tolerance of 1e-9 of the box size for points computed on a face
  bool inside(const V3<double>& p) const {
    const double v[3] = {p.x, p.y, p.z};
    for (int d = 0; d < 3; ++d) {
      const double tol = 1e-9 * (hi[d] - lo[d]);
      if (!(v[d] >= lo[d] - tol && v[d] <= hi[d] + tol)) return false;
    }
    return true;
  }
};

// findIntersectVolume(mesh, l, l0): squared distance from l0 to the nearest
// intersection of the line l0 + s l with a face plane that lies in the
// volume (Inf if none), and that intersection.
double find_intersect(const Cuboid& cb, const V3<double>& l, const V3<double>& l0,
                      V3<double>& node) {
  double dmin = std::numeric_limits<double>::infinity();
  node = mk(0.0, 0.0, 0.0);
  for (int i = 0; i < 6; ++i) {
    const double old = dmin;
    const double den = dot(l, cb.n(i));
    const double num = dot(cb.m(i) - l0, cb.n(i));
    if (den == 0) continue;  // parallel: no point, or the whole line (Inf)
    const double dd = num / den;
    const V3<double> x = l0 + dd * l;
    if (std::isinf(x.x) || std::isinf(x.y) || std::isinf(x.z)) continue;
    if (!cb.inside(x)) continue;
    const V3<double> r = x - l0;
    dmin = std::min((r.x * r.x + r.y * r.y) + r.z * r.z, old);
    if (dmin < old) node = x;
  }
  return dmin;
}

}  // namespace

extern "C" {

// calc_uTilde: u (3 x P, point by point) += displacement of every segment
// of the network at the points.  k2 = (omNuInv8pi, om2NuOmNuInv8pi);
// label: nodeTypeDln per node (Int32); vertices 3 x 8, faceNorm and
// faceMidPt 3 x 6 of the cuboid mesh.
void orc_u_tilde(const double* k2, int64_t numSeg, const int64_t* links, const double* coord,
                 const double* bVec, const int32_t* label, int64_t numNodeCols,
                 const double* vertices, const double* faceNorm, const double* faceMidPt,
                 const double* pts, int64_t P, double* u) {
  const UT k{k2[0], k2[1]};
  Cuboid cb{faceNorm, faceMidPt, {INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}};
  V3<double> C = mk(0.0, 0.0, 0.0);
  for (int v = 0; v < 8; ++v) {
    C = C + mk(vertices[3 * v], vertices[3 * v + 1], vertices[3 * v + 2]);
    for (int d = 0; d < 3; ++d) {
      cb.lo[d] = std::min(cb.lo[d], vertices[3 * v + d]);
      cb.hi[d] = std::max(cb.hi[d], vertices[3 * v + d]);
    }
  }
  C = C / 8.0;
  const int32_t srfMob = 3, srfFix = 4, ext = 5;
  std::vector<int64_t> surfCon((size_t)numNodeCols + 1, 0);
  auto X = [&](int64_t node) {
    return mk(coord[3 * (node - 1)], coord[3 * (node - 1) + 1], coord[3 * (node - 1) + 2]);
  };
  for (int64_t i = 0; i < numSeg; ++i) {
    const int64_t n1 = links[2 * i], n2 = links[2 * i + 1];
    const int32_t l1 = label[n1 - 1], l2 = label[n2 - 1];
    const bool s1 = l1 == srfMob || l1 == srfFix, s2 = l2 == srfMob || l2 == srfFix;
    if (s1 && l2 == ext) { surfCon[n2] = n1; continue; }
    if (s2 && l1 == ext) { surfCon[n1] = n2; continue; }
    const V3<double> A = X(n1), B = X(n2);
    const V3<double> b = mk(bVec[3 * i], bVec[3 * i + 1], bVec[3 * i + 2]);
    if (l1 != ext && l2 != ext) {
      tri_full(k, A, B, C, b, pts, P, u);
      continue;
    }
    // External segment: project both ends back to the surface along the
    // normals of faces 1, 3, 5 (a cuboid needs only those).
    const double inf = std::numeric_limits<double>::infinity();
    double dA = inf, dB = inf, dC = inf;
    V3<double> iA = mk(inf, inf, inf), iB = iA;
    for (int j = 0; j < 6; j += 2) {
      V3<double> xa, xb;
      const double ta = find_intersect(cb, cb.n(j), A, xa);
      const double tb = find_intersect(cb, cb.n(j), B, xb);
      if (ta < dA) { dA = ta; iA = xa; }
      if (tb < dB) { dB = tb; iB = xb; }
    }
    // No intersection (projected out of an edge or corner): walk towards
    // the centroid over the face planes, as the reference does (its plane
    // distance adds the face midpoint, and one minimum serves both ends).
    if (std::isinf(dA)) {
      V3<double> t = A;
      for (int f = 0; f < 6; ++f) {
        const double dist = sq(dot(t + cb.m(f), cb.n(f)));
        const V3<double> x = t - cb.n(f) * std::sqrt(dist);
        const double dc = norm(x - C);
        if (dc <= dC) { dC = dc; t = x; }
      }
      iA = t;
    }
    if (std::isinf(dB)) {
      V3<double> t = B;
      for (int f = 0; f < 6; ++f) {
        const double dist = sq(dot(t + cb.m(f), cb.n(f)));
        const V3<double> x = t - cb.n(f) * std::sqrt(dist);
        const double dc = norm(x - C);
        if (dc <= dC) { dC = dc; t = x; }
      }
      iB = t;
    }
    V3<double> Ap = iA, Bp = iB;
    if (surfCon[n1] != 0) Ap = X(surfCon[n1]);
    if (surfCon[n2] != 0) Bp = X(surfCon[n2]);
    const V3<double> Cp = (A + Bp) / 2.0;
    tri_solid(Ap, Bp, C, b, pts, P, u);
    tri_solid(Ap, A, Cp, b, pts, P, u);
    tri_solid(A, B, Cp, b, pts, P, u);
    tri_solid(B, Bp, Cp, b, pts, P, u);
    tri_solid(Bp, Ap, Cp, b, pts, P, u);
  }
}

}  // extern "C"

// ---- detectCollision ------------------------------------------------------
// minimumDistance (Misc.jl:314-367) and detectCollision
// (DislocationTopologyCollision.jl:1-222).
namespace {

struct MinDist {
  double distSq, dDistSqDt, L1, L2;
};

MinDist minimum_distance(const V3<double>& x0, const V3<double>& x1, const V3<double>& y0,
                         const V3<double>& y1, const V3<double>& vx0, const V3<double>& vx1,
                         const V3<double>& vy0, const V3<double>& vy1) {
  const double eps = std::numeric_limits<double>::epsilon();
  const V3<double> s1 = x1 - x0, s2 = y1 - y0, v1 = vx1 - vx0, v2 = vy1 - vy0;
  const double A = dot(s1, s1);
  const double B = dot(2.0 * s1, x0 - y0);
  const double C = dot(2.0 * s1, s2);
  const double D = dot(2.0 * s2, y0 - x0);
  const double E = dot(s2, s2);
  const double G = C * C - 4.0 * A * E;
  double L1, L2;
  if (A < eps) {
    L1 = 0;
    L2 = E < eps ? 0.0 : -D / E / 2.0;
  } else if (E < eps) {
    L2 = 0;
    L1 = A < eps ? 0.0 : -B / A / 2.0;
  } else if (std::fabs(G) < eps) {
    const double d[4] = {dot(y0 - x0, y0 - x0), dot(y1 - x0, y1 - x0), dot(y0 - x1, y0 - x1),
                         dot(y1 - x1, y1 - x1)};
    int idx = 1;  // findmin: first minimum, 1-based
    for (int q = 1; q < 4; ++q)
      if (d[q] < d[idx - 1]) idx = q + 1;
    L1 = std::floor(idx / 2.0);
    L2 = (double)((idx - 1) % 2);
  } else {
    L2 = (2.0 * A * D + B * C) / G;
    L1 = (C * L2 - B) / A / 2.0;
  }
  L1 = L1 < 0 ? 0.0 : (L1 > 1 ? 1.0 : L1);
  L2 = L2 < 0 ? 0.0 : (L2 > 1 ? 1.0 : L2);
  const V3<double> r = ((x0 + s1 * L1) - y0) - s2 * L2;
  const V3<double> vr = ((vx0 + v1 * L1) - vy0) - v2 * L2;
  return MinDist{dot(r, r), 2.0 * dot(vr, r), L1, L2};
}

}  // namespace

extern "C" {

void orc_minimum_distance(const double* x0, const double* x1, const double* y0, const double* y1,
                          const double* vx0, const double* vx1, const double* vy0,
                          const double* vy1, double* out) {
  auto v = [](const double* p) { return mk(p[0], p[1], p[2]); };
  const MinDist m = minimum_distance(v(x0), v(x1), v(y0), v(y1), v(vx0), v(vx1), v(vy0), v(vy1));
  out[0] = m.distSq;
  out[1] = m.dDistSqDt;
  out[2] = m.L1;
  out[3] = m.L2;
}

// detectCollision: res[0..7] = collision (0/1), type (0 none, 1 hinge,
// 2 twoLine), n1s1, n2s1, n1s2, n2s2, s1, s2; resL = (L1, L2).  skip: nSkip
// pairs of link numbers.  label: nodeTypeDln per node (Int32).
void orc_detect_collision(double collisionDistSq, int64_t numNode, int64_t numSeg,
                          const int64_t* links, const int64_t* connectivity, int64_t ldConn,
                          const double* coord, const double* nodeVel, const double* bVec,
                          const int32_t* label, const int64_t* skip, int64_t nSkip, int64_t* res,
                          double* resL) {
  const double eps = std::numeric_limits<double>::epsilon();
  auto X = [&](int64_t n) { return mk(coord[3 * (n - 1)], coord[3 * (n - 1) + 1], coord[3 * (n - 1) + 2]); };
  auto Vn = [&](int64_t n) {
    return mk(nodeVel[3 * (n - 1)], nodeVel[3 * (n - 1) + 1], nodeVel[3 * (n - 1) + 2]);
  };
  auto mobile = [&](int64_t n) { return label[n - 1] == 1 || label[n - 1] == 3; };
  auto skipped = [&](int64_t a, int64_t b) {
    for (int64_t q = 0; q < nSkip; ++q) {
      const int64_t s1 = skip[2 * q], s2 = skip[2 * q + 1];
      if ((a == s1 && b == s2) || (a == s2 && b == s1)) return true;
    }
    return false;
  };
  double best = 0, L1 = 0, L2 = 0;
  bool collision = false;
  int64_t type = 0, n1s1 = 0, n2s1 = 0, n1s2 = 0, n2s2 = 0, s1 = 0, s2 = 0;
  // hinges
  for (int64_t i = 1; i <= numNode; ++i) {
    if (!mobile(i)) continue;
    const int64_t* cn = connectivity + ldConn * (i - 1);
    for (int64_t j = 1; j <= cn[0]; ++j) {
      // DislocationTopologyCollision.jl:30-41: link1, connectedNode1 and
      // tVecN1 belong to the j loop; the swap at :52-68 assigns them, so it
      // carries over to later k (link 1 is the longest of j..k-1).
      int64_t lk1 = cn[2 * j - 1];
      int64_t c1 = links[2 * (lk1 - 1) + (2 - cn[2 * j])];
      const V3<double> xi = X(i);
      const V3<double> t1 = xi - X(c1);
      double tn1 = dot(t1, t1);
      for (int64_t k = j + 1; k <= cn[0]; ++k) {
        int64_t lk2 = cn[2 * k - 1];
        int64_t c2 = links[2 * (lk2 - 1) + (2 - cn[2 * k])];
        const V3<double> t2 = xi - X(c2);
        double tn2 = dot(t2, t2);
        if (tn1 < tn2) {  // :52-68, collapse the shorter one into a point
          std::swap(lk1, lk2);
          std::swap(c1, c2);
          std::swap(tn1, tn2);
        }
        if (skipped(lk1, lk2)) continue;
        const MinDist m = minimum_distance(xi, X(c1), X(c2), X(c2), Vn(i), Vn(c1), Vn(c2), Vn(c2));
        L1 = m.L1;
        L2 = m.L2;
        if (m.distSq < collisionDistSq && m.dDistSqDt < -eps) {
          const double key = 1 / m.distSq;
          if (key > best) {
            best = key;
            collision = true;
            type = 1;
            n1s1 = i;
            n2s1 = c1;
            n1s2 = c2;
            n2s2 = c2;
            s1 = lk1;
            s2 = lk2;
          }
        }
      }
    }
  }
  if (!collision) {
    for (int64_t i = 1; i <= numSeg; ++i) {
      const int64_t a1 = links[2 * (i - 1)], a2 = links[2 * (i - 1) + 1];
      if (!mobile(a1) || !mobile(a2)) continue;
      const V3<double> x0 = X(a1), x1 = X(a2);
      const V3<double> bi = mk(bVec[3 * (i - 1)], bVec[3 * (i - 1) + 1], bVec[3 * (i - 1) + 2]);
      for (int64_t j = i + 1; j <= numSeg; ++j) {
        const int64_t b1 = links[2 * (j - 1)], b2 = links[2 * (j - 1) + 1];
        if (a1 == b1 || a1 == b2 || a2 == b1 || a2 == b2) continue;
        if (!mobile(b1) || !mobile(b2)) continue;
        const V3<double> y0 = X(b1), y1 = X(b2);
        const V3<double> bj = mk(bVec[3 * (j - 1)], bVec[3 * (j - 1) + 1], bVec[3 * (j - 1) + 2]);
        const double par = dot(x1 - x0, y1 - y0);
        if ((dot(bi - bj, bi - bj) < eps && par > eps) || (dot(bi + bj, bi + bj) < eps && par < eps))
          continue;
        if (skipped(i, j)) continue;
        const MinDist m = minimum_distance(x0, x1, y0, y1, Vn(a1), Vn(a2), Vn(b1), Vn(b2));
        L1 = m.L1;
        L2 = m.L2;
        if (m.distSq < collisionDistSq && m.dDistSqDt < -eps) {
          const double key = 1 / m.distSq;
          if (key > best) {
            best = key;
            collision = true;
            type = 2;
            n1s1 = a1;
            n2s1 = a2;
            n1s2 = b1;
            n2s2 = b2;
            s1 = i;
            s2 = j;
          }
        }
      }
    }
  }
  const int64_t r[8] = {collision ? 1 : 0, type, n1s1, n2s1, n1s2, n2s2, s1, s2};
  for (int q = 0; q < 8; ++q) res[q] = r[q];
  resL[0] = L1;
  resL[1] = L2;
}

}  // extern "C"
