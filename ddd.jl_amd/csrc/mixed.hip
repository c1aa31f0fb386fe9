// The mixed-precision all-pairs kernels (the two-phase tile kernel and the
// near-parallel passes of mode 1, from tiles.hpp).  This translation unit is compiled with
// floating-point contraction and fast FP32 division and square root: the
// mixed path is an approximation held to the FP32 quadrature's accuracy
// (~1e-6 per node against the FP64 path), and its classification of pairs
// uses explicitly rounded arithmetic, so it matches the count and listing
// passes bit for bit.
#include "kernels.hpp"
#include "pair_special.hpp"
#include "tiles.hpp"

namespace ddd {

namespace {
__global__ void k_zero_tiles(unsigned long long* c) { *c = 0; }

// Mixed-precision pair tiles (no reference counterpart), persistent: the
// workgroups take tiles t0 + counters[0] ... from an atomic counter and do
// each in two phases, so that neither phase carries the other's registers:
//  1. the 4,096 slots in (step, lane) order: well-separated pairs
//     (midpoint distance > farDist and |t1 x t2| > sinMin) run inline in
//     FP32 by quadrature (far_side); the other generic pairs go, in slot
//     order, to the workgroup's staging area (4,096 uint16 in HBM);
//  2. the staged pairs in FP64 (FAST corner integrals: an approximation,
//     like the FP32 quadrature), 64 at a time (near_batch), added in staging
//     order.
// Near-parallel pairs were listed and evaluated beforehand, as in the plain
// kernel.  The number of FP64 pairs (near-parallel ones included) is added
// to counters[1].
__global__ __launch_bounds__(kTile, 2) void k_segseg_tiles_mixed(const double* __restrict__ seg,
                                                                 int64_t ld, int n, int64_t tile0,
                                                                 int64_t tile1, SegK<double> K,
                                                                 SigK<float> KF, double farDist2,
                                                                 double sinMin2,
                                                                 double* __restrict__ rowPart,
                                                                 double* __restrict__ colPart,
                                                                 const long long* __restrict__ specOff,
                                                                 const unsigned short* __restrict__ specSlot,
                                                                 const double* __restrict__ specRes,
                                                                 int64_t specTotal, int64_t specBase,
                                                                 unsigned short* __restrict__ stage,
                                                                 unsigned long long* __restrict__ counters) {
  const int lane = threadIdx.x;
  __shared__ double si[kTileFields][kTile];
  __shared__ double sj[kTileFields][kTile];
  __shared__ double cacc[6][kTile];
  __shared__ double sres[3][kTile];
  __shared__ unsigned long long rowM[kTile], colM[kTile];
  __shared__ long long tNext;
  unsigned short* const q = stage + (size_t)blockIdx.x * (kTile * kTile);
  const uint64_t ltMask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  unsigned exact = 0;

  for (;;) {
    if (lane == 0) tNext = tile0 + (long long)atomicAdd(&counters[0], 1ull);
    __syncthreads();
    const int64_t t = tNext;
    __syncthreads();
    if (t >= tile1) break;
    int I, J;
    tile_blocks(t, I, J);
    bool rowExt;
    uint64_t colExt;
    stage_blocks(seg, ld, I, J, lane, si, sj, rowExt, colExt);
#pragma unroll
    for (int c = 0; c < 6; ++c) cacc[c][lane] = 0.0;
    __syncthreads();

    const Seg<double> A = seg_from_lds(si, lane);
    const int64_t i = (int64_t)I * kTile + lane;
    const bool rowLive = (i < n) && !rowExt;
    const bool diag = (I == J);
    const v3<double> midA = mid_rn(A.n1, A.n2);
    v3<double> a1 = V(0.0, 0.0, 0.0), a2 = V(0.0, 0.0, 0.0);
    int nq = 0;  // wave-uniform
    for (int k = 0; k < kTile; ++k) {
      const int s = (lane + k) & 63;
      const int64_t j = (int64_t)J * kTile + s;
      const bool live = rowLive && (j < n) && !((colExt >> s) & 1ull) && (!diag || s > lane);
      const Seg<double> B = seg_from_lds(sj, s);
      double c;
      const double omcSq = one_minus_c2(A.t, B.t, c);
      const bool special = is_special<1>(live, omcSq, midA, B.n1, B.n2, farDist2, sinMin2);
      const bool far = is_far(omcSq, midA, B.n1, B.n2, farDist2, sinMin2);
      const bool queued = live && !far && !special;
      exact += __popcll(__ballot(special));
      if (live && far) {
        v3<double> F1, F2, F3, F4;
        far_side<float>(KF, A, B, F1, F2);
        far_side<float>(KF, B, A, F3, F4);
        a1 = a1 + F1;
        a2 = a2 + F2;
        cacc[0][s] += F3.x; cacc[1][s] += F3.y; cacc[2][s] += F3.z;
        cacc[3][s] += F4.x; cacc[4][s] += F4.y; cacc[5][s] += F4.z;
      }
      const uint64_t m = __ballot(queued);
      if (queued) q[nq + __popcll(m & ltMask)] = (unsigned short)((k << 6) | lane);
      nq += __popcll(m);
      __syncthreads();
    }
    exact += (unsigned)nq;
    __syncthreads();  // the staged entries are visible to every lane
    for (int b = 0; b < nq; b += kTile) {
      const int cnt = min(nq - b, kTile);
      const unsigned e = lane < cnt ? q[b + lane] : 0u;
      near_batch(e, cnt, lane,
                 [&](int rl, int sl, v3<double>& F1, v3<double>& F2, v3<double>& F3, v3<double>& F4) {
                   const Seg<double> P = seg_from_lds(si, rl), Q = seg_from_lds(sj, sl);
                   double c;
                   const double omc = one_minus_c2(P.t, Q.t, c);
                   pair_generic<double, true, true, true>(K, P, Q, c, omc, F1, F2, F3, F4);
                 },
                 sres, cacc, rowM, colM, a1, a2);
    }
    tile_finish(t, lane, a1, a2, cacc, rowPart, colPart, specOff, specSlot, specRes, specTotal,
                tile0, specBase);
  }
  if (lane == 0 && exact) atomicAdd(&counters[1], (unsigned long long)exact);
}
}  // namespace

void launch_special_pass_mixed(const double* seg, int64_t ld, int n, const MixedRule& mr,
                               unsigned* cnt, const SpecialList* sp, int64_t t0, int64_t t1,
                               hipStream_t st) {
  const int64_t maxGrid = 1ll << 30;
  const double fd2 = mr.farDist2(), sm2 = mr.sinMin2();
  for (int64_t a = t0; a < t1; a += maxGrid) {
    const int64_t count = (t1 - a < maxGrid) ? (t1 - a) : maxGrid;
    if (sp)
      hipLaunchKernelGGL((k_special_pass<1, 1>), dim3((unsigned)count), dim3(kTile), 0, st, seg, ld, n,
                         a, fd2, sm2, cnt, sp->off, sp->base, sp->slot, sp->tile);
    else
      hipLaunchKernelGGL((k_special_pass<1, 0>), dim3((unsigned)count), dim3(kTile), 0, st, seg, ld, n,
                         a, fd2, sm2, cnt, nullptr, (int64_t)0, nullptr, nullptr);
  }
}

void launch_segseg_tiles_mixed(const double* seg, int64_t ld, int n, const SegK<double>& K,
                               double* rowPart, double* colPart, const SpecialList& sp,
                               const TileRange& r, hipStream_t st, const MixedRule& mr,
                               unsigned short* stage, unsigned long long* counters) {
  hipLaunchKernelGGL(k_zero_tiles, dim3(1), dim3(1), 0, st, counters);
  const int64_t blocks = std::min<int64_t>(r.t1() - r.t0(), kMixedBlocks);
  hipLaunchKernelGGL(k_segseg_tiles_mixed, dim3((unsigned)blocks), dim3(kTile), 0, st, seg, ld, n,
                     r.t0(), r.t1(), K, mr.K, mr.farDist2(), mr.sinMin2(), rowPart, colPart, sp.off,
                     sp.slot, sp.res, sp.total, sp.base, stage, counters);
}

}  // namespace ddd
