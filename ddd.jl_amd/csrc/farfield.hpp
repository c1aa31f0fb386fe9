// Hybrid near/far field: cell grid, moments, far-field operator, launches.
#pragma once
#include "kernels.hpp"
#include "pair_special.hpp"

#include <vector>

namespace ddd {

constexpr int kMoments = 90;  // cell features: orders 0, 1, 2 of b (x) t L
constexpr int kPoly = 10;     // local basis: 1, x, y, z, x^2, y^2, z^2, xy, xz, yz
constexpr int kLocal = 6 * kPoly;          // stress components x basis
constexpr int kChan = kLocal * kMoments;   // entries of one W(o)
// Real channels travel through the FFTs two per complex channel.
constexpr int kMomPairs = kMoments / 2;
constexpr int kLocPairs = kLocal / 2;
constexpr int kChanPairs = kChan / 2;

// Cubic symmetry of the far-field operator.  A signed permutation R of the
// axes, (R v)[perm[k]] = sgn[k] v[k], carries a source configuration, its
// offset, moments and local expansion into another valid one; isotropic
// elasticity therefore gives W(R o) = det(R) T_L(R) W(o) T_M(R)^T, with T_M,
// T_L the signed index permutations below (a reflection, det = -1, also
// reverses the right-handed Burgers circuit, which flips the stress).  Every
// offset o is R o' of a canonical offset o' (components sorted by magnitude,
// descending, non-negative).
struct AxisMap {
  int perm[3];
  int sgn[3];
};

// Index of the symmetric pair (i, j) in the order xx, yy, zz, xy, xz, yz (the
// stress components and the quadratic terms of the local basis).
__host__ __device__ inline int sym_pair(int i, int j) {
  if (i == j) return i;
  const int a = i < j ? i : j, b = i < j ? j : i;
  return a == 0 ? (b == 1 ? 3 : 4) : 5;
}
__host__ __device__ inline void pair_axes(int p, int& i, int& j) {
  i = p < 3 ? p : (p == 5 ? 1 : 0);
  j = p < 3 ? p : (p == 3 ? 1 : 2);
}

// o = R o' with o' canonical; ties keep the lower axis first.
__host__ __device__ inline void canonical_offset(int ox, int oy, int oz, int oc[3], AxisMap& R) {
  const int o[3] = {ox, oy, oz};
  const int m[3] = {ox < 0 ? -ox : ox, oy < 0 ? -oy : oy, oz < 0 ? -oz : oz};
  int ix[3] = {0, 1, 2};
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2 - a; ++b)
      if (m[ix[b + 1]] > m[ix[b]]) {
        const int t = ix[b];
        ix[b] = ix[b + 1];
        ix[b + 1] = t;
      }
  for (int k = 0; k < 3; ++k) {
    oc[k] = m[ix[k]];
    R.perm[k] = ix[k];
    R.sgn[k] = o[ix[k]] < 0 ? -1 : 1;
  }
}

// +1 for a rotation, -1 for a reflection.
__host__ __device__ inline int axis_det(const AxisMap& R) {
  const int inv = (R.perm[0] > R.perm[1]) + (R.perm[0] > R.perm[2]) + (R.perm[1] > R.perm[2]);
  return ((inv & 1) ? -1 : 1) * R.sgn[0] * R.sgn[1] * R.sgn[2];
}

// Tetrahedral index of a canonical offset (a >= b >= c >= 0).
__host__ __device__ inline int canonical_index(int a, int b, int c) {
  return a * (a + 1) * (a + 2) / 6 + b * (b + 1) / 2 + c;
}

// Moment k (layout of the moments kernel: 3a + j; 9 + 9a + 3j + m;
// 36 + 18a + 6j + p) under R: (T_M v)[k2] = s v[k].
__host__ __device__ inline void moment_map(const AxisMap& R, int k, int& k2, int& s) {
  if (k < 9) {
    const int a = k / 3, j = k % 3;
    k2 = 3 * R.perm[a] + R.perm[j];
    s = R.sgn[a] * R.sgn[j];
  } else if (k < 36) {
    const int e = k - 9, a = e / 9, j = (e / 3) % 3, m = e % 3;
    k2 = 9 + 9 * R.perm[a] + 3 * R.perm[j] + R.perm[m];
    s = R.sgn[a] * R.sgn[j] * R.sgn[m];
  } else {
    const int e = k - 36, a = e / 18, j = (e / 6) % 3, p = e % 6;
    int u, v;
    pair_axes(p, u, v);
    k2 = 36 + 18 * R.perm[a] + 6 * R.perm[j] + sym_pair(R.perm[u], R.perm[v]);
    s = R.sgn[a] * R.sgn[j] * R.sgn[u] * R.sgn[v];
  }
}

// Local-expansion entry q = 10 c + l (stress component c, basis term l of
// 1, x, y, z, xx, yy, zz, xy, xz, yz) under R: (T_L v)[q2] = s v[q].
__host__ __device__ inline void local_map(const AxisMap& R, int q, int& q2, int& s) {
  const int c = q / 10, l = q % 10;
  int i, j;
  pair_axes(c, i, j);
  const int c2 = sym_pair(R.perm[i], R.perm[j]);
  s = R.sgn[i] * R.sgn[j];
  int l2 = 0;
  if (l >= 1 && l <= 3) {
    l2 = 1 + R.perm[l - 1];
    s *= R.sgn[l - 1];
  } else if (l >= 4) {
    int u, v;
    pair_axes(l - 4, u, v);
    l2 = 4 + sym_pair(R.perm[u], R.perm[v]);
    s *= R.sgn[u] * R.sgn[v];
  }
  q2 = 10 * c2 + l2;
}

// R then S (apply S first): (R S v)[R.perm[S.perm[k]]] = R.sgn[S.perm[k]] S.sgn[k] v[k].
__host__ __device__ inline AxisMap compose(const AxisMap& R, const AxisMap& S) {
  AxisMap C;
  for (int k = 0; k < 3; ++k) {
    C.perm[k] = R.perm[S.perm[k]];
    C.sgn[k] = R.sgn[S.perm[k]] * S.sgn[k];
  }
  return C;
}
__host__ __device__ inline AxisMap inverse(const AxisMap& R) {
  AxisMap I;
  for (int k = 0; k < 3; ++k) {
    I.perm[R.perm[k]] = k;
    I.sgn[R.perm[k]] = R.sgn[k];
  }
  return I;
}
// The k-th of the 48 signed permutations (k = 8 p + signs).
__host__ __device__ inline AxisMap axis_map(int k) {
  const int P[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  AxisMap R;
  for (int e = 0; e < 3; ++e) {
    R.perm[e] = P[k / 8][e];
    R.sgn[e] = ((k >> e) & 1) ? -1 : 1;
  }
  return R;
}

struct GridDev {
  double lo[3];
  double h;
  int G;
};

int fft_size_for(int G);
void launch_cells(const double* seg, int64_t ld, int n, const GridDev& g, int* cellOf,
                  unsigned* counts, unsigned* cursor, int* start, int* tmp, int* order,
                  hipStream_t st);
void launch_moments(const double* seg, int64_t ld, const GridDev& g, const int* start,
                    const int* order, float* M, hipStream_t st);
void launch_far_field(const float* M, const GridDev& g, const float2* What, const float2* tw,
                      float2* Mc, float2* Lc, float* L, hipStream_t st);
// What: [half frequency][90 moments][60 locals] complex, f_z in [0, NP/2] (NP^2 (NP/2 + 1)
// frequencies); scratch: kChanPairs NP^3 complex; Mc: kMomPairs NP^3;
// Lc: kLocPairs NP^3.
void launch_operator_spectrum(const float* W, const GridDev& g, int rn, const float2* tw,
                              float2* scratch, float2* What, hipStream_t st);
void launch_far_apply(const double* seg, int64_t ld, int n, const GridDev& g, const int* cellOf,
                      const float* L, double* out, hipStream_t st);
// Device arrays of the near-field tile list (see farfield.hip).
struct NearListDev {
  // inputs: cell binning
  const int* cellOf;
  const unsigned* counts;
  const int* start;
  const int* order;
  const int* mrank;     // Morton rank of each cell
  const int* rankCell;  // cell of each Morton rank
  // outputs
  unsigned* mcount;
  int* mstart;          // nc + 1
  int64_t ldS;          // pad64(numSeg): sorted copy leading dimension
  double* segS;         // F_COUNT x ldS
  int* key;             // packed cell coordinates per sorted segment
  int* perm;            // original index per sorted segment (-1: padding)
  int* box;             // 6 per block
  unsigned* rowCount;
  int* rowStart;        // nbS + 1; rowStart[nbS] = number of tiles
  unsigned* colCount;
  int* colStart;        // nbS + 1
  int2* list;           // (I, J) per tile
  int* tileIdx;         // nbS x nbS, J >= I
  int* colList;
  int rn;
};

void morton_ranks(int G, std::vector<int>& mrank, std::vector<int>& rankCell);
// Sorted copy, block boxes, per-row tile counts and rowStart (its last entry
// is the number of tiles); then, with list and colList sized, the lists.
void launch_near_list_count(const double* seg, int64_t ld, int n, const GridDev& g, int rn,
                            NearListDev& L, hipStream_t st);
void launch_near_list_fill(const NearListDev& L, hipStream_t st);
void launch_near_reduce(const NearListDev& L, int nc, const double* rowPart, const double* colPart,
                        double* out, hipStream_t st);
// Near-field tiles: workgroups of the main kernel and the staging entries
// (uint16) each one owns (two per slot of a tile); stage holds
// kNearGrid * kNearStage entries.  The main kernel lists the near-parallel
// pairs (spList, capacity spCap; specOff / specCnt per tile) and counts them
// in counters[3]; counters (four) are zeroed before the call.
constexpr int kNearGrid = 2048;
constexpr int kNearStage = 2 * kTile * kTile;
void launch_near_tiles(const double* segS, int64_t ldS, const int* key, int rn, const int2* list,
                       int nTiles, const SegK<double>& K, const MixedRule& mr, double* rowPart,
                       double* colPart, unsigned short* stage, unsigned long long* spList,
                       long long spCap, long long* specOff, int* specCnt,
                       unsigned long long* counters, hipStream_t st);
// The listed near-parallel pairs (total <= spCap): evaluated into res
// (12 x spCap doubles; fix: spCap entries and a counter, the pairs that need
// the branch's corrective sub-pairs) and added to their tiles' partial sums.
void launch_near_special(const double* segS, int64_t ldS, const int2* list, int nTiles,
                         const SegK<double>& K, const unsigned long long* spList, long long total,
                         long long spCap, const long long* specOff, const int* specCnt, double* res,
                         long long* fix, unsigned long long* fixCount, double* rowPart,
                         double* colPart, unsigned* flags, hipStream_t st);

}  // namespace ddd
