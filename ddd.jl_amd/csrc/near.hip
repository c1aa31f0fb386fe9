// Near-field tile kernels of the hybrid path (no reference counterpart).
//
// This file is compiled with floating-point contraction enabled and its FP64
// pair takes the FAST corner integrals of pair_math.hpp: the hybrid is an
// approximation measured against the FP64 path (its far field is good to
// ~1e-3), so its near field is not held to the FP64 path's operation-for-
// operation agreement with the oracle.  Fused multiply-adds and refined
// hardware reciprocals move a near pair's forces within the pair formula's
// own conditioning (up to ~3e-7 per node against the IEEE oracle) and remove
// many of the FP64 pair's instructions.
#include <algorithm>

#include "batch.hpp"
#include "farfield.hpp"
#include "pair_special.hpp"

namespace ddd {

namespace {

// Segment fields staged in LDS: n1, n2, b, t, |t|.  The external flag is
// folded into the cell key (-1) instead.
constexpr int kLdsFields = F_TN + 1;

__device__ __forceinline__ Seg<double> seg_from_lds(const double (*s)[kTile], int l) {
  Seg<double> r;
  r.n1 = V(s[F_N1X][l], s[F_N1Y][l], s[F_N1Z][l]);
  r.n2 = V(s[F_N2X][l], s[F_N2Y][l], s[F_N2Z][l]);
  r.b = V(s[F_BX][l], s[F_BY][l], s[F_BZ][l]);
  r.t = V(s[F_TX][l], s[F_TY][l], s[F_TZ][l]);
  r.tN = s[F_TN][l];
  return r;
}

// Tile staging: segments of row block I and column block J into LDS, cell
// keys (-1 for padding) into ki (row, register) and kj (columns, LDS).
__device__ __forceinline__ void stage_tile(const double* __restrict__ seg, int64_t ld,
                                           const int* __restrict__ key, int I, int J, int lane,
                                           double (*si)[kTile], double (*sj)[kTile], int* kj,
                                           int& ki) {
  const int64_t i = (int64_t)I * kTile + lane;
  const int64_t jl = (int64_t)J * kTile + lane;
#pragma unroll
  for (int f = 0; f < kLdsFields; ++f) {
    si[f][lane] = seg[f * ld + i];
    sj[f][lane] = seg[f * ld + jl];
  }
  ki = seg[F_EXT * ld + i] == 0.0 ? key[i] : -1;
  kj[lane] = seg[F_EXT * ld + jl] == 0.0 ? key[jl] : -1;
}

// Pair classes of the near field.
enum : int { kNone = 0, kFP32 = 1, kFP64 = 2, kPar = 3 };

// Class of the pair (row lane, column slot s) of a staged tile: not near
// (cells farther than rn in some axis, padding, or not above the diagonal of
// a diagonal tile), well separated and non-parallel (kFP32: midpoints
// farther than farDist, sin^2 of the angle above sinMin^2), near-parallel
// (kPar: 1 - c^2 <= sqrt(eps), the reference's special branch) or generic
// (kFP64).  Every rounding is explicit, so the decisions do not depend on
// how the compiler contracts this translation unit.
__device__ __forceinline__ int near_class(const double (*si)[kTile], const double (*sj)[kTile],
                                          const double (*cmid)[kTile], int ki, int kb, int lane,
                                          int s, bool diag, int rn, double farDist2,
                                          double sinMin2) {
  // Without early returns: every lane evaluates the tests and the class is
  // selected, so the slot loop has no exec-mask branches.
  const int dCell = max(max(abs((ki & 1023) - (kb & 1023)),
                            abs(((ki >> 10) & 1023) - ((kb >> 10) & 1023))),
                        abs((ki >> 20) - (kb >> 20)));
  const bool live = (ki >= 0) & (kb >= 0) & !(diag & (s <= lane)) & (dCell <= rn);
  double c;
  const double omcSq = one_minus_c2(V(si[F_TX][lane], si[F_TY][lane], si[F_TZ][lane]),
                                    V(sj[F_TX][s], sj[F_TY][s], sj[F_TZ][s]), c);
  // twice the midpoint separation (cmid: n1 + n2 of the columns)
  const double dx = __dsub_rn(cmid[0][s], __dadd_rn(si[F_N1X][lane], si[F_N2X][lane]));
  const double dy = __dsub_rn(cmid[1][s], __dadd_rn(si[F_N1Y][lane], si[F_N2Y][lane]));
  const double dz = __dsub_rn(cmid[2][s], __dadd_rn(si[F_N1Z][lane], si[F_N2Z][lane]));
  const double d2x4 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
  const int cls = ((d2x4 > 4.0 * farDist2) & (omcSq > sinMin2)) ? kFP32
                  : (omcSq > DDD_SQRT_EPS ? kFP64 : kPar);
  return live ? cls : kNone;
}

// Near-field tiles of the hybrid path: tile t = (I, J) of the cell-sorted
// copy, from a list; persistent workgroups take tiles from counters[2].
// Only pairs whose cells lie within Chebyshev distance rn take part (cell
// coordinates packed in key).  A tile is done in two phases:
//  1. classification of its 4,096 slots (near_class) into this workgroup's
//     staging area in HBM, in slot order: FP64 entries at the front of the
//     first 4,096 entries, FP32 entries at its back, near-parallel entries
//     in the second 4,096.  The near-parallel entries are then copied to
//     the global list spList at an offset taken from counters[3] (tile in
//     the high bits, slot code in the low 16), with specOff[t] / specCnt[t]
//     recording where; entries past spCap are counted but not written (the
//     caller enlarges the list and runs again);
//  2. the FP64 entries (the generic analytic pair), then the FP32 entries
//     (FP32 quadrature of both sides, far_side), 64 at a time (near_batch2).
// Sums are in a fixed order, so results are reproducible.  Keeping the
// classification state out of the evaluation, column totals in LDS and
// the near-parallel branch out of the kernel, the ~200-register FP64 pair
// fits two wavefronts per SIMD; the LDS is 20 KB for eight workgroups per
// CU (the batch buffer doubles as classification storage, the mask arrays
// as the tile and list-base broadcasts).
// counters[0] += near pairs, counters[1] += FP64 pairs (generic and
// near-parallel).
__global__ __launch_bounds__(kTile, 2) void k_near_tiles(const double* __restrict__ seg, int64_t ld,
                                                         const int* __restrict__ key, int rn,
                                                         const int2* __restrict__ list, int nTiles,
                                                         SegK<double> K, SigK<float> KF,
                                                         double farDist2, double sinMin2,
                                                         double* __restrict__ rowPart,
                                                         double* __restrict__ colPart,
                                                         unsigned short* __restrict__ stage,
                                                         unsigned long long* __restrict__ spList,
                                                         long long spCap,
                                                         long long* __restrict__ specOff,
                                                         int* __restrict__ specCnt,
                                                         unsigned long long* __restrict__ counters) {
  const int lane = threadIdx.x;

  __shared__ double si[kLdsFields][kTile];
  __shared__ double sj[kLdsFields][kTile];
  __shared__ double cacc[6][kTile];  // column totals of the current tile
  // Two forces of each pair of a batch (near_batch2); during the
  // classification, rows 0-2 hold the columns' n1 + n2 and row 3 their
  // cell keys (kj).
  __shared__ double sres[6][kTile];
  // Ownership masks of a batch; outside the batches, rowM[0] passes the
  // next tile's index and colM[0] the tile's base in the near-parallel list.
  // (20 KB of LDS in all: eight workgroups a CU.)
  __shared__ unsigned long long rowM[kTile], colM[kTile];
  int* const kj = reinterpret_cast<int*>(&sres[3][0]);
  int& tNext = *reinterpret_cast<int*>(&rowM[0]);
  long long& spBase = *reinterpret_cast<long long*>(&colM[0]);

  constexpr int kSlots = kTile * kTile;
  unsigned short* const q = stage + (size_t)blockIdx.x * kNearStage;
  unsigned nNear = 0, nExact = 0;
  const uint64_t ltMask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

  for (;;) {
    if (lane == 0) tNext = (int)atomicAdd(&counters[2], 1ull);
    __syncthreads();
    const int t = tNext;
    __syncthreads();
    if (t >= nTiles) break;
    const int I = list[t].x, J = list[t].y;
    int ki;
    stage_tile(seg, ld, key, I, J, lane, si, sj, kj, ki);
#pragma unroll
    for (int c = 0; c < 6; ++c) cacc[c][lane] = 0.0;
    // n1 + n2 of the columns for the classification, in the batch result
    // buffer (free until the evaluation)
    sres[0][lane] = __dadd_rn(sj[F_N1X][lane], sj[F_N2X][lane]);
    sres[1][lane] = __dadd_rn(sj[F_N1Y][lane], sj[F_N2Y][lane]);
    sres[2][lane] = __dadd_rn(sj[F_N1Z][lane], sj[F_N2Z][lane]);
    __syncthreads();

    // 1. classification
    const bool diag = (I == J);
    int n64 = 0, n32 = 0, nsp = 0;
    for (int k = 0; k < kTile; ++k) {
      const int s = (lane + k) & 63;
      const int cls = near_class(si, sj, sres, ki, kj[s], lane, s, diag, rn, farDist2, sinMin2);
      const uint64_t w32 = __ballot(cls == kFP32), w64 = __ballot(cls == kFP64);
      const uint64_t wsp = __ballot(cls == kPar);
      const unsigned short code = (unsigned short)((k << 6) | lane);
      // one store a slot: FP64 entries from the front, FP32 from the back of
      // the first kSlots entries, near-parallel ones in the second kSlots
      const bool f32 = cls == kFP32, f64 = cls == kFP64;
      const uint64_t wsel = f32 ? w32 : (f64 ? w64 : wsp);
      const int pos = (f32 ? n32 : (f64 ? n64 : nsp)) + __popcll(wsel & ltMask);
      const int dest = f32 ? kSlots - 1 - pos : (f64 ? pos : kSlots + pos);
      if (cls != kNone) q[dest] = code;
      n64 += __popcll(w64);
      n32 += __popcll(w32);
      nsp += __popcll(wsp);
    }
    nNear += (unsigned)(n64 + n32 + nsp);
    nExact += (unsigned)(n64 + nsp);
    if (lane == 0) {
      const long long base = nsp ? (long long)atomicAdd(&counters[3], (unsigned long long)nsp) : 0;
      spBase = base;
      specOff[t] = base;
      specCnt[t] = nsp;
    }
    __syncthreads();  // the staged entries and spBase are visible to every lane
    for (int b = lane; b < nsp; b += kTile)
      if (spBase + b < spCap)
        spList[spBase + b] = ((unsigned long long)t << 16) | q[kSlots + b];

    // 2. evaluation
    v3<double> a1 = V(0.0, 0.0, 0.0), a2 = V(0.0, 0.0, 0.0);  // row forces (segment I*64+lane)
    for (int b = 0; b < n64; b += kTile) {
      const int cnt = min(n64 - b, kTile);
      const unsigned e = lane < cnt ? q[b + lane] : 0u;
      near_batch2(e, cnt, lane,
                 [&](int rl, int sl, v3<double>& F1, v3<double>& F2, v3<double>& F3, v3<double>& F4) {
                   const Seg<double> P = seg_from_lds(si, rl), Q = seg_from_lds(sj, sl);
                   double c;
                   const double omc = one_minus_c2(P.t, Q.t, c);
                   pair_generic<double, true, true, true>(K, P, Q, c, omc, F1, F2, F3, F4);
                 },
                 sres, cacc, rowM, colM, a1, a2);
    }
    for (int b = 0; b < n32; b += kTile) {
      const int cnt = min(n32 - b, kTile);
      const unsigned e = lane < cnt ? q[kSlots - 1 - (b + lane)] : 0u;
      near_batch2(e, cnt, lane,
                 [&](int rl, int sl, v3<double>& F1, v3<double>& F2, v3<double>& F3, v3<double>& F4) {
                   const Seg<double> P = seg_from_lds(si, rl), Q = seg_from_lds(sj, sl);
                   far_side<float>(KF, P, Q, F1, F2);
                   far_side<float>(KF, Q, P, F3, F4);
                 },
                 sres, cacc, rowM, colM, a1, a2);
    }

    double* rp = rowPart + (int64_t)t * 6 * kTile;
    double* cp = colPart + (int64_t)t * 6 * kTile;
    rp[0 * kTile + lane] = a1.x; rp[1 * kTile + lane] = a1.y; rp[2 * kTile + lane] = a1.z;
    rp[3 * kTile + lane] = a2.x; rp[4 * kTile + lane] = a2.y; rp[5 * kTile + lane] = a2.z;
#pragma unroll
    for (int c = 0; c < 6; ++c) cp[c * kTile + lane] = cacc[c][lane];
  }
  if (lane == 0 && nNear) {
    atomicAdd(&counters[0], (unsigned long long)nNear);
    atomicAdd(&counters[1], (unsigned long long)nExact);
  }
}

// Near-parallel pairs of the near-field tiles (the reference's special
// branch, calcParSegSegForce), one list entry per lane: pass A of the branch,
// forces F1..F4 into res[12][spCap]; the pairs that need corrective
// sub-pairs are listed in fix (claimed by wave ballots: order immaterial,
// each pair has its own result slot).
__global__ __launch_bounds__(kTile) void k_near_special_eval(const double* __restrict__ seg, int64_t ld,
                                                             const int2* __restrict__ list,
                                                             const unsigned long long* __restrict__ spList,
                                                             long long total, long long spCap,
                                                             SegK<double> K, double* __restrict__ res,
                                                             long long* __restrict__ fix,
                                                             unsigned long long* __restrict__ fixCount) {
  const long long e = (long long)blockIdx.x * kTile + threadIdx.x;
  bool need = false;
  if (e < total) {
    const unsigned long long en = spList[e];
    const int t = (int)(en >> 16);
    const unsigned code = (unsigned)(en & 0xffffu);
    const int rl = code & 63, sl = (rl + (code >> 6)) & 63;
    const int2 IJ = list[t];
    const Seg<double> P = seg_from(seg, ld, (int64_t)IJ.x * kTile + rl);
    const Seg<double> Q = seg_from(seg, ld, (int64_t)IJ.y * kTile + sl);
    v3<double> F1, F2, F3, F4;
    need = special_pair_base(K, P, Q, F1, F2, F3, F4);
    const double v[12] = {F1.x, F1.y, F1.z, F2.x, F2.y, F2.z, F3.x, F3.y, F3.z, F4.x, F4.y, F4.z};
#pragma unroll
    for (int c = 0; c < 12; ++c) res[c * spCap + e] = v[c];
  }
  const uint64_t m = __ballot(need);
  if (m == 0) return;
  const int lane = threadIdx.x;
  unsigned long long base = 0;
  if (lane == __builtin_ctzll(m)) base = atomicAdd(fixCount, (unsigned long long)__popcll(m));
  base = __shfl(base, __builtin_ctzll(m), kTile);
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  if (need) fix[base + __popcll(m & lt)] = e;
}

__global__ void k_near_zero_count(unsigned long long* c) { *c = 0; }

// Pass B for the listed pairs: their corrective sub-pairs, added to their
// forces in res (grid-stride over the list, whose length is on the device).
__global__ __launch_bounds__(kTile) void k_near_special_fix(const double* __restrict__ seg, int64_t ld,
                                                            const int2* __restrict__ list,
                                                            const unsigned long long* __restrict__ spList,
                                                            long long spCap, SegK<double> K,
                                                            double* __restrict__ res,
                                                            const long long* __restrict__ fix,
                                                            const unsigned long long* __restrict__ fixCount,
                                                            unsigned* __restrict__ flags) {
  const long long cnt = (long long)*fixCount;
  for (long long q0 = (long long)blockIdx.x * kTile + threadIdx.x; q0 < cnt;
       q0 += (long long)gridDim.x * kTile) {
    const long long e = fix[q0];
    const unsigned long long en = spList[e];
    const int t = (int)(en >> 16);
    const unsigned code = (unsigned)(en & 0xffffu);
    const int rl = code & 63, sl = (rl + (code >> 6)) & 63;
    const int2 IJ = list[t];
    const Seg<double> P = seg_from(seg, ld, (int64_t)IJ.x * kTile + rl);
    const Seg<double> Q = seg_from(seg, ld, (int64_t)IJ.y * kTile + sl);
    v3<double> F1 = V(res[0 * spCap + e], res[1 * spCap + e], res[2 * spCap + e]);
    v3<double> F2 = V(res[3 * spCap + e], res[4 * spCap + e], res[5 * spCap + e]);
    v3<double> F3 = V(res[6 * spCap + e], res[7 * spCap + e], res[8 * spCap + e]);
    v3<double> F4 = V(res[9 * spCap + e], res[10 * spCap + e], res[11 * spCap + e]);
    special_pair_fix(K, P, Q, F1, F2, F3, F4, flags);
    const double v[12] = {F1.x, F1.y, F1.z, F2.x, F2.y, F2.z, F3.x, F3.y, F3.z, F4.x, F4.y, F4.z};
#pragma unroll
    for (int c = 0; c < 12; ++c) res[c * spCap + e] = v[c];
  }
}

// Adds the near-parallel forces of tile t (blockIdx.x) to its partial sums,
// 64 entries at a time: each lane takes, in slot order, the entries whose
// row (column) slot it is, marked with LDS atomic ORs.
__global__ __launch_bounds__(kTile) void k_near_special_add(const unsigned long long* __restrict__ spList,
                                                            const double* __restrict__ res,
                                                            long long spCap,
                                                            const long long* __restrict__ specOff,
                                                            const int* __restrict__ specCnt,
                                                            double* __restrict__ rowPart,
                                                            double* __restrict__ colPart) {
  __shared__ double sv[12][kTile];
  __shared__ unsigned long long rowM[kTile], colM[kTile];
  const int t = blockIdx.x, lane = threadIdx.x;
  const int n = specCnt[t];
  if (n == 0) return;
  const long long off = specOff[t];
  double a[6] = {0, 0, 0, 0, 0, 0}, b[6] = {0, 0, 0, 0, 0, 0};
  for (int e0 = 0; e0 < n; e0 += kTile) {
    const int cnt = min(n - e0, kTile);
    rowM[lane] = 0;
    colM[lane] = 0;
    __syncthreads();
    if (lane < cnt) {
      const long long e = off + e0 + lane;
      const unsigned code = (unsigned)(spList[e] & 0xffffu);
      const int rl = code & 63, sl = (rl + (code >> 6)) & 63;
      atomicOr(&rowM[rl], 1ull << lane);
      atomicOr(&colM[sl], 1ull << lane);
#pragma unroll
      for (int c = 0; c < 12; ++c) sv[c][lane] = res[c * spCap + e];
    }
    __syncthreads();
    for (uint64_t m = rowM[lane]; m; m &= m - 1) {
      const int e = __builtin_ctzll(m);
#pragma unroll
      for (int c = 0; c < 6; ++c) a[c] += sv[c][e];
    }
    for (uint64_t m = colM[lane]; m; m &= m - 1) {
      const int e = __builtin_ctzll(m);
#pragma unroll
      for (int c = 0; c < 6; ++c) b[c] += sv[6 + c][e];
    }
    __syncthreads();
  }
  double* rp = rowPart + (int64_t)t * 6 * kTile + lane;
  double* cp = colPart + (int64_t)t * 6 * kTile + lane;
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    rp[c * kTile] += a[c];
    cp[c * kTile] += b[c];
  }
}

}  // namespace

void launch_near_tiles(const double* segS, int64_t ldS, const int* key, int rn, const int2* list,
                       int nTiles, const SegK<double>& K, const MixedRule& mr, double* rowPart,
                       double* colPart, unsigned short* stage, unsigned long long* spList,
                       long long spCap, long long* specOff, int* specCnt,
                       unsigned long long* counters, hipStream_t st) {
  if (nTiles <= 0) return;
  // Persistent workgroups take tiles from counters[2] (zeroed by the
  // caller).  kNearGrid workgroups, two per SIMD of the 256 CUs, keep every
  // SIMD busy until the list ends; each owns kNearStage staging entries.
  const int grid = nTiles < kNearGrid ? nTiles : kNearGrid;
  hipLaunchKernelGGL(k_near_tiles, dim3((unsigned)grid), dim3(kTile), 0, st, segS, ldS, key, rn, list,
                     nTiles, K, mr.K, mr.farDist2(), mr.sinMin2(), rowPart,
                     colPart, stage, spList, spCap, specOff, specCnt, counters);
}

void launch_near_special(const double* segS, int64_t ldS, const int2* list, int nTiles,
                         const SegK<double>& K, const unsigned long long* spList, long long total,
                         long long spCap, const long long* specOff, const int* specCnt, double* res,
                         long long* fix, unsigned long long* fixCount, double* rowPart,
                         double* colPart, unsigned* flags, hipStream_t st) {
  if (nTiles <= 0 || total <= 0) return;
  hipLaunchKernelGGL(k_near_zero_count, dim3(1), dim3(1), 0, st, fixCount);
  hipLaunchKernelGGL(k_near_special_eval, dim3((unsigned)((total + kTile - 1) / kTile)), dim3(kTile), 0, st,
                     segS, ldS, list, spList, total, spCap, K, res, fix, fixCount);
  const long long blocks = std::min<long long>((total + kTile - 1) / kTile, 2048);
  hipLaunchKernelGGL(k_near_special_fix, dim3((unsigned)blocks), dim3(kTile), 0, st, segS, ldS, list,
                     spList, spCap, K, res, fix, fixCount, flags);
  hipLaunchKernelGGL(k_near_special_add, dim3((unsigned)nTiles), dim3(kTile), 0, st, spList, res, spCap,
                     specOff, specCnt, rowPart, colPart);
}

}  // namespace ddd
