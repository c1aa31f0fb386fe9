// Near-field tile kernel of the hybrid path (no reference counterpart).
//
// This file is compiled with floating-point contraction enabled: the hybrid
// is an approximation measured against the FP64 path (its far field is good
// to ~1e-3), so its near field is not held to the FP64 path's operation-for-
// operation agreement with the oracle.  Fused multiply-adds move a near
// pair's forces within the pair formula's own conditioning (up to 6e-8 per
// node against the uncontracted oracle at separations of ~500) and remove
// many of the FP64 pair's instructions.
#include "farfield.hpp"
#include "pair_special.hpp"

namespace ddd {

namespace {

__device__ __forceinline__ Seg<double> seg_from_lds(const double (*s)[kTile], int l) {
  Seg<double> r;
  r.n1 = V(s[F_N1X][l], s[F_N1Y][l], s[F_N1Z][l]);
  r.n2 = V(s[F_N2X][l], s[F_N2Y][l], s[F_N2Z][l]);
  r.b = V(s[F_BX][l], s[F_BY][l], s[F_BZ][l]);
  r.t = V(s[F_TX][l], s[F_TY][l], s[F_TZ][l]);
  r.tN = s[F_TN][l];
  return r;
}

constexpr int kQueue = 2 * kTile;

// Near-field tiles of the hybrid path: tile t = (I, J) of the cell-sorted
// copy, from a list.  Only pairs whose cells lie within Chebyshev distance rn
// take part (cell coordinates packed in key).  Each slot is classified and
// compacted into one of three queues -- well-separated, non-parallel pairs
// (FP32 quadrature of both sides, far_side), the other generic pairs (FP64
// analytic) and near-parallel pairs (FP64, the reference's special branch)
// -- and each queue is evaluated 64 pairs at a time, one per lane.  When a
// pair enters one of the first two queues, its row lane and its column lane
// (row + k mod 64, told by a shuffle) mark the queue position in 128-bit
// masks; after an evaluation each lane adds the results at its own marked
// positions in increasing order, so sums are deterministic and row and
// column totals stay in registers.
// counters[0] += near pairs, counters[1] += FP64 pairs.
struct QMask {
  uint64_t rowLo = 0, rowHi = 0, colLo = 0, colHi = 0;
};

__device__ __forceinline__ void mark(uint64_t& lo, uint64_t& hi, int pos) {
  if (pos < 64) lo |= 1ull << pos;
  else hi |= 1ull << (pos - 64);
}

__global__ __launch_bounds__(kTile) void k_near_tiles(const double* __restrict__ seg, int64_t ld,
                                                      const int* __restrict__ key, int rn,
                                                      const int2* __restrict__ list, int nTiles,
                                                      SegK<double> K, SigK<float> KF,
                                                      double farDist2, double sinMin2,
                                                      double* __restrict__ rowPart,
                                                      double* __restrict__ colPart,
                                                      unsigned* __restrict__ flags,
                                                      unsigned long long* __restrict__ counters) {
  const int lane = threadIdx.x;

  __shared__ double si[F_COUNT][kTile];
  __shared__ double sj[F_COUNT][kTile];
  __shared__ int kj[kTile];
  __shared__ double sres[12][kTile];
  __shared__ unsigned short q64[kQueue];
  __shared__ unsigned short q32[kQueue];
  __shared__ unsigned short qsp[kQueue];
  __shared__ unsigned short qcE[kQueue];  // carried near-parallel pairs: slot ...
  __shared__ int qcT[kQueue];             // ... and tile
  __shared__ int kR[kTile], kC[kTile];
  __shared__ int tNext;

  unsigned nNear = 0, nExact = 0;
  int ncar = 0;  // carried near-parallel pairs (wave-uniform)
  const uint64_t ltMask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

  // The first cnt carried near-parallel pairs, evaluated one per lane with
  // their segments read from HBM, added to the partial sums their tiles have
  // already written.  Entries with the same target row (or column) are added
  // in queue order: in round r each entry whose target has r earlier entries
  // in the batch adds its result, so no two lanes touch one address in a
  // round and the sums do not depend on how the batch was cut.
  auto flushCarried = [&](int cnt) {
    __threadfence();  // the tiles' partial sums are written
    int t = 0, rl = 0, sl = 0;
    if (lane < cnt) {
      t = qcT[lane];
      const unsigned e = qcE[lane];
      rl = e & 63;
      sl = (rl + (e >> 6)) & 63;
      const int2 IJ = list[t];
      const Seg<double> P = seg_from(seg, ld, (int64_t)IJ.x * kTile + rl);
      const Seg<double> Q = seg_from(seg, ld, (int64_t)IJ.y * kTile + sl);
      v3<double> F1, F2, F3, F4;
      special_pair_dev(K, P, Q, F1, F2, F3, F4, flags);
      sres[0][lane] = F1.x; sres[1][lane] = F1.y; sres[2][lane] = F1.z;
      sres[3][lane] = F2.x; sres[4][lane] = F2.y; sres[5][lane] = F2.z;
      sres[6][lane] = F3.x; sres[7][lane] = F3.y; sres[8][lane] = F3.z;
      sres[9][lane] = F4.x; sres[10][lane] = F4.y; sres[11][lane] = F4.z;
      kR[lane] = t * kTile + rl;
      kC[lane] = t * kTile + sl;
    }
    __syncthreads();
    int rankR = 0, rankC = 0;
    if (lane < cnt)
      for (int e = 0; e < lane; ++e) {
        rankR += kR[e] == kR[lane];
        rankC += kC[e] == kC[lane];
      }
    int top = lane < cnt ? (rankR > rankC ? rankR : rankC) : 0;
    for (int off = 32; off >= 1; off >>= 1) {
      const int o = __shfl_xor(top, off, kTile);
      top = o > top ? o : top;
    }
    for (int r = 0; r <= top; ++r) {
      if (lane < cnt && rankR == r) {
        double* p = rowPart + (int64_t)t * 6 * kTile + rl;
#pragma unroll
        for (int c = 0; c < 6; ++c) p[c * kTile] += sres[c][lane];
      }
      if (lane < cnt && rankC == r) {
        double* p = colPart + (int64_t)t * 6 * kTile + sl;
#pragma unroll
        for (int c = 0; c < 6; ++c) p[c * kTile] += sres[6 + c][lane];
      }
      __threadfence();  // this round's stores before the next round's loads
    }
    // drop the flushed entries
    const int rest = ncar - cnt;
    const unsigned short ke = (lane < rest) ? qcE[cnt + lane] : 0;
    const int kt = (lane < rest) ? qcT[cnt + lane] : 0;
    __syncthreads();
    if (lane < rest) {
      qcE[lane] = ke;
      qcT[lane] = kt;
    }
    __syncthreads();
    ncar = rest;
  };

  for (;;) {
    if (lane == 0) tNext = (int)atomicAdd(&counters[2], 1ull);
    __syncthreads();
    const int t = tNext;
    __syncthreads();
    if (t >= nTiles) break;
    const int I = list[t].x, J = list[t].y;

    const int64_t i = (int64_t)I * kTile + lane;
    const int64_t jl = (int64_t)J * kTile + lane;
#pragma unroll
    for (int f = 0; f < F_COUNT; ++f) {
      si[f][lane] = seg[f * ld + i];
      sj[f][lane] = seg[f * ld + jl];
    }
    const int ki = key[i];
    kj[lane] = key[jl];
    __syncthreads();

    const Seg<double> A = seg_from_lds(si, lane);
    const bool rowLive = si[F_EXT][lane] == 0.0;
    const bool diag = (I == J);
    const int ax = ki & 1023, ay = (ki >> 10) & 1023, az = ki >> 20;
    v3<double> a1 = V(0.0, 0.0, 0.0), a2 = V(0.0, 0.0, 0.0);  // row forces (segment i)
    v3<double> b1 = V(0.0, 0.0, 0.0), b2 = V(0.0, 0.0, 0.0);  // column forces (segment J*64+lane)
    int n64 = 0, n32 = 0, nsp = 0;
    QMask m64, m32;
    const v3<double> midA = V(0.5 * (A.n1.x + A.n2.x), 0.5 * (A.n1.y + A.n2.y), 0.5 * (A.n1.z + A.n2.z));

    // Evaluate the first cnt (<= 64) entries of queue q, add this lane's
    // marked results, and drop those entries from the masks.
    auto flush = [&](unsigned short* q, int& nq, QMask& mk, bool exact) {
      const int cnt = nq < kTile ? nq : kTile;
      if (lane < cnt) {
        const unsigned e = q[lane];
        const int rl = e & 63, sl = (rl + (e >> 6)) & 63;
        const Seg<double> P = seg_from_lds(si, rl), Q = seg_from_lds(sj, sl);
        v3<double> F1, F2, F3, F4;
        if (exact) {
          double c;
          const double omc = one_minus_c2(P.t, Q.t, c);
          pair_generic<double, true, true>(K, P, Q, c, omc, F1, F2, F3, F4);
        } else {
          far_side<float>(KF, P, Q, F1, F2);
          far_side<float>(KF, Q, P, F3, F4);
        }
        sres[0][lane] = F1.x; sres[1][lane] = F1.y; sres[2][lane] = F1.z;
        sres[3][lane] = F2.x; sres[4][lane] = F2.y; sres[5][lane] = F2.z;
        sres[6][lane] = F3.x; sres[7][lane] = F3.y; sres[8][lane] = F3.z;
        sres[9][lane] = F4.x; sres[10][lane] = F4.y; sres[11][lane] = F4.z;
      }
      __syncthreads();
      for (uint64_t r = mk.rowLo; r; r &= r - 1) {  // positions < cnt: the whole low word
        const int e = __builtin_ctzll(r);
        a1 = a1 + V(sres[0][e], sres[1][e], sres[2][e]);
        a2 = a2 + V(sres[3][e], sres[4][e], sres[5][e]);
      }
      for (uint64_t r = mk.colLo; r; r &= r - 1) {
        const int e = __builtin_ctzll(r);
        b1 = b1 + V(sres[6][e], sres[7][e], sres[8][e]);
        b2 = b2 + V(sres[9][e], sres[10][e], sres[11][e]);
      }
      // entries 64.. move to 0..
      const int rest = nq - cnt;
      const unsigned short keep = (lane < rest) ? q[kTile + lane] : 0;
      __syncthreads();
      if (lane < rest) q[lane] = keep;
      __syncthreads();
      mk.rowLo = mk.rowHi; mk.rowHi = 0;
      mk.colLo = mk.colHi; mk.colHi = 0;
      nq = rest;
    };

    // 64 near-parallel pairs of this tile (their queue filled within the
    // tile): evaluated from LDS and added in queue order by a plain scan.
    auto flushSpecial = [&]() {
      const unsigned e = qsp[lane];
      const int rl = e & 63, sl = (rl + (e >> 6)) & 63;
      const Seg<double> P = seg_from_lds(si, rl), Q = seg_from_lds(sj, sl);
      v3<double> F1, F2, F3, F4;
      special_pair_dev(K, P, Q, F1, F2, F3, F4, flags);
      sres[0][lane] = F1.x; sres[1][lane] = F1.y; sres[2][lane] = F1.z;
      sres[3][lane] = F2.x; sres[4][lane] = F2.y; sres[5][lane] = F2.z;
      sres[6][lane] = F3.x; sres[7][lane] = F3.y; sres[8][lane] = F3.z;
      sres[9][lane] = F4.x; sres[10][lane] = F4.y; sres[11][lane] = F4.z;
      __syncthreads();
      for (int q = 0; q < kTile; ++q) {
        const unsigned qe = qsp[q];
        const int rq = qe & 63, sq = (rq + (qe >> 6)) & 63;
        if (rq == lane) {
          a1 = a1 + V(sres[0][q], sres[1][q], sres[2][q]);
          a2 = a2 + V(sres[3][q], sres[4][q], sres[5][q]);
        }
        if (sq == lane) {
          b1 = b1 + V(sres[6][q], sres[7][q], sres[8][q]);
          b2 = b2 + V(sres[9][q], sres[10][q], sres[11][q]);
        }
      }
      const int rest = nsp - kTile;
      const unsigned short keep = (lane < rest) ? qsp[kTile + lane] : 0;
      __syncthreads();
      if (lane < rest) qsp[lane] = keep;
      __syncthreads();
      nsp = rest;
    };

    for (int k = 0; k < kTile; ++k) {
      const int s = (lane + k) & 63;
      const int kb = kj[s];
      const bool cellNear = abs(ax - (kb & 1023)) <= rn && abs(ay - ((kb >> 10) & 1023)) <= rn &&
                            abs(az - (kb >> 20)) <= rn;
      const bool live = rowLive && cellNear && sj[F_EXT][s] == 0.0 && (!diag || s > lane);
      const v3<double> tB = V(sj[F_TX][s], sj[F_TY][s], sj[F_TZ][s]);
      double c;
      const double omcSq = one_minus_c2(A.t, tB, c);
      const double dx = 0.5 * (sj[F_N1X][s] + sj[F_N2X][s]) - midA.x,
                   dy = 0.5 * (sj[F_N1Y][s] + sj[F_N2Y][s]) - midA.y,
                   dz = 0.5 * (sj[F_N1Z][s] + sj[F_N2Z][s]) - midA.z;
      const bool far = (dx * dx + dy * dy + dz * dz > farDist2) && (omcSq > sinMin2);
      const bool is32 = live && far, is64 = live && !far && omcSq > DDD_SQRT_EPS;
      const bool isSp = live && !far && !(omcSq > DDD_SQRT_EPS);
      const uint64_t w32 = __ballot(is32), w64 = __ballot(is64), wsp = __ballot(isSp);
      if (isSp) qsp[nsp + __popcll(wsp & ltMask)] = (unsigned short)((k << 6) | lane);
      const int p32 = is32 ? n32 + __popcll(w32 & ltMask) : -1;
      const int p64 = is64 ? n64 + __popcll(w64 & ltMask) : -1;
      if (is32) {
        q32[p32] = (unsigned short)((k << 6) | lane);
        mark(m32.rowLo, m32.rowHi, p32);
      }
      if (is64) {
        q64[p64] = (unsigned short)((k << 6) | lane);
        mark(m64.rowLo, m64.rowHi, p64);
      }
      // this lane is the column of the row lane (lane - k) mod 64
      const int src = (lane - k) & 63;
      const int c32 = __shfl(p32, src, kTile), c64 = __shfl(p64, src, kTile);
      if (c32 >= 0) mark(m32.colLo, m32.colHi, c32);
      if (c64 >= 0) mark(m64.colLo, m64.colHi, c64);
      n32 += __popcll(w32);
      n64 += __popcll(w64);
      nsp += __popcll(wsp);
      nNear += __popcll(w32) + __popcll(w64) + __popcll(wsp);
      nExact += __popcll(w64) + __popcll(wsp);
      __syncthreads();
      if (n32 >= kTile) flush(q32, n32, m32, false);
      if (n64 >= kTile) flush(q64, n64, m64, true);
      if (nsp >= kTile) flushSpecial();
    }
    if (n32 > 0) flush(q32, n32, m32, false);
    if (n64 > 0) flush(q64, n64, m64, true);

    double* rp = rowPart + (int64_t)t * 6 * kTile;
    double* cp = colPart + (int64_t)t * 6 * kTile;
    rp[0 * kTile + lane] = a1.x; rp[1 * kTile + lane] = a1.y; rp[2 * kTile + lane] = a1.z;
    rp[3 * kTile + lane] = a2.x; rp[4 * kTile + lane] = a2.y; rp[5 * kTile + lane] = a2.z;
    cp[0 * kTile + lane] = b1.x; cp[1 * kTile + lane] = b1.y; cp[2 * kTile + lane] = b1.z;
    cp[3 * kTile + lane] = b2.x; cp[4 * kTile + lane] = b2.y; cp[5 * kTile + lane] = b2.z;

    // This tile's remaining near-parallel pairs join the carried queue,
    // which is evaluated whenever it holds a full wavefront's worth.
    if (lane < nsp) {
      qcE[ncar + lane] = qsp[lane];
      qcT[ncar + lane] = t;
    }
    ncar += nsp;
    __syncthreads();
    if (ncar >= kTile) flushCarried(kTile);
  }
  if (ncar > 0) flushCarried(ncar);
  if (lane == 0 && nNear) {
    atomicAdd(&counters[0], (unsigned long long)nNear);
    atomicAdd(&counters[1], (unsigned long long)nExact);
  }
}

}  // namespace

void launch_near_tiles(const double* segS, int64_t ldS, const int* key, int rn, const int2* list,
                       int nTiles, const SegK<double>& K, const MixedRule& mr, double* rowPart,
                       double* colPart, unsigned* flags, unsigned long long* counters,
                       hipStream_t st) {
  if (nTiles <= 0) return;
  // Persistent workgroups take tiles from counters[2] (zeroed by the caller);
  // two per SIMD of the 256 CUs keep every SIMD busy until the list ends.
  const int grid = nTiles < 2048 ? nTiles : 2048;
  hipLaunchKernelGGL(k_near_tiles, dim3((unsigned)grid), dim3(kTile), 0, st, segS, ldS, key, rn, list,
                     nTiles, K, mr.K, mr.farDist * mr.farDist, mr.sinMin * mr.sinMin, rowPart,
                     colPart, flags, counters);
}

}  // namespace ddd
