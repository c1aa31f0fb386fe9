// The all-pairs tile kernels and their near-parallel passes, as templates on
// the mode (0: plain FP64, 1: mixed precision).  segseg.hip instantiates the
// plain kernels with the FP64 path's strict rounding; mixed.hip the mixed
// ones, compiled with contraction and fast FP32 division (an approximation
// held to the FP32 quadrature's own accuracy).  Everything sits in an
// unnamed namespace, so each translation unit has its own copy.
#pragma once
#include "batch.hpp"
#include "kernels.hpp"
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

// Tile t of the upper triangle, enumerated column-block major:
// t = J (J + 1) / 2 + I with 0 <= I <= J.
__device__ __forceinline__ void tile_blocks(int64_t t, int& I, int& J) {
  J = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((int64_t)J * (J + 1) / 2 > t) --J;
  while ((int64_t)(J + 1) * (J + 2) / 2 <= t) ++J;
  I = (int)(t - (int64_t)J * (J + 1) / 2);
}

// Midpoint of a segment and the mixed rule's far test in explicitly rounded
// arithmetic, so that every kernel classifies a pair identically whether or
// not its translation unit contracts floating-point expressions.
__device__ __forceinline__ v3<double> mid_rn(v3<double> a, v3<double> b) {
  return V(__dmul_rn(0.5, __dadd_rn(a.x, b.x)), __dmul_rn(0.5, __dadd_rn(a.y, b.y)),
           __dmul_rn(0.5, __dadd_rn(a.z, b.z)));
}
__device__ __forceinline__ bool is_far(double omcSq, v3<double> midA, v3<double> n1B,
                                       v3<double> n2B, double farDist2, double sinMin2) {
  const v3<double> mB = mid_rn(n1B, n2B);
  const double dx = __dsub_rn(mB.x, midA.x), dy = __dsub_rn(mB.y, midA.y),
               dz = __dsub_rn(mB.z, midA.z);
  const double d2 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
  return d2 > farDist2 && omcSq > sinMin2;
}

// The classification both the count pass and the tile kernel use, so that
// they agree bit for bit: a live pair is near-parallel (the reference's
// special branch, 1 - c^2 <= sqrt(eps)) unless, in the mixed mode, it is a
// well-separated pair for the FP32 quadrature.
template <int MODE>
__device__ __forceinline__ bool is_special(bool live, double omcSq, v3<double> midA,
                                           v3<double> n1B, v3<double> n2B, double farDist2,
                                           double sinMin2) {
  if (!live || omcSq > DDD_SQRT_EPS) return false;
  if (MODE == 0) return true;
  return !is_far(omcSq, midA, n1B, n2B, farDist2, sinMin2);
}

// Near-parallel pairs per tile, one wavefront per tile in the tile kernel's
// slot order (k, lane): LIST = 0 counts them (cnt[t]); LIST = 1 writes them
// to the tile's range off[t] .. off[t + 1] of the pair list, at position
// e - base of the chunk's list.
template <int MODE, int LIST>
__global__ __launch_bounds__(kTile) void k_special_pass(const double* __restrict__ seg, int64_t ld,
                                                        int n, int64_t tile0, double farDist2,
                                                        double sinMin2, unsigned* __restrict__ cnt,
                                                        const long long* __restrict__ off,
                                                        int64_t base,
                                                        unsigned short* __restrict__ slot,
                                                        int* __restrict__ tileOf) {
  const int64_t t = tile0 + blockIdx.x;
  int I, J;
  tile_blocks(t, I, J);
  const int lane = threadIdx.x;
  // staged column fields: t, ext, and (mixed mode, for is_far) n1, n2
  constexpr int NF = MODE == 0 ? 4 : 10;
  __shared__ double sj[NF][kTile];
  const int64_t i = (int64_t)I * kTile + lane;
  const int64_t jl = (int64_t)J * kTile + lane;
  const int fl[10] = {F_TX, F_TY, F_TZ, F_EXT, F_N1X, F_N1Y, F_N1Z, F_N2X, F_N2Y, F_N2Z};
#pragma unroll
  for (int f = 0; f < NF; ++f) sj[f][lane] = seg[fl[f] * ld + jl];
  __syncthreads();
  const v3<double> tA = V(seg[F_TX * ld + i], seg[F_TY * ld + i], seg[F_TZ * ld + i]);
  v3<double> midA = V(0.0, 0.0, 0.0);
  if constexpr (MODE != 0)
    midA = mid_rn(V(seg[F_N1X * ld + i], seg[F_N1Y * ld + i], seg[F_N1Z * ld + i]),
                  V(seg[F_N2X * ld + i], seg[F_N2Y * ld + i], seg[F_N2Z * ld + i]));
  const bool rowLive = (i < n) && seg[F_EXT * ld + i] == 0.0;
  const bool diag = (I == J);
  const uint64_t ltMask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  long long c = LIST ? off[t] : 0;
  const long long cEnd = LIST ? off[t + 1] : 0;
  for (int k = 0; k < kTile; ++k) {
    const int s = (lane + k) & 63;
    const int64_t j = (int64_t)J * kTile + s;
    const bool live = rowLive && (j < n) && sj[3][s] == 0.0 && (!diag || s > lane);
    double cc;
    const double omcSq = one_minus_c2(tA, V(sj[0][s], sj[1][s], sj[2][s]), cc);
    v3<double> n1B = V(0.0, 0.0, 0.0), n2B = n1B;
    if constexpr (MODE != 0) {
      n1B = V(sj[4][s], sj[5][s], sj[6][s]);
      n2B = V(sj[7][s], sj[8][s], sj[9][s]);
    }
    const bool sp = is_special<MODE>(live, omcSq, midA, n1B, n2B, farDist2, sinMin2);
    const uint64_t m = __ballot(sp);
    if (LIST && sp) {
      const long long pos = c + __popcll(m & ltMask);
      if (pos < cEnd) {  // always true: both passes classify identically
        slot[pos - base] = (unsigned short)((k << 6) | lane);
        tileOf[pos - base] = (int)t;
      }
    }
    c += __popcll(m);
  }
  if (!LIST && lane == 0) cnt[t] = (unsigned)c;
}

// Staged fields of the tile kernels: n1, n2, b, t, 1/L (external flags are
// kept as ballot masks in registers).  With at most 20 KiB of LDS a
// workgroup, eight fit on a CU (two wavefronts per SIMD).
constexpr int kTileFields = F_EXT;

__device__ __forceinline__ void stage_blocks(const double* __restrict__ seg, int64_t ld, int I,
                                             int J, int lane, double (*si)[kTile],
                                             double (*sj)[kTile], bool& rowExt,
                                             uint64_t& colExt) {
  const int64_t i = (int64_t)I * kTile + lane;
  const int64_t jl = (int64_t)J * kTile + lane;
#pragma unroll
  for (int f = 0; f < kTileFields; ++f) {
    si[f][lane] = seg[f * ld + i];
    sj[f][lane] = seg[f * ld + jl];
  }
  rowExt = seg[F_EXT * ld + i] != 0.0;
  colExt = __ballot(seg[F_EXT * ld + jl] != 0.0);
}

// Adds tile t's near-parallel pairs, evaluated beforehand, in list order,
// and writes the tile's partial sums (slot t - partBase).
__device__ __forceinline__ void tile_finish(int64_t t, int lane, v3<double> a1, v3<double> a2,
                                            double (*cacc)[kTile],
                                            double* __restrict__ rowPart,
                                            double* __restrict__ colPart,
                                            const long long* __restrict__ specOff,
                                            const unsigned short* __restrict__ specSlot,
                                            const double* __restrict__ specRes, int64_t specTotal,
                                            int64_t partBase, int64_t specBase) {
  for (long long e = specOff[t] - specBase, e1 = specOff[t + 1] - specBase; e < e1; ++e) {
    const unsigned q = specSlot[e];
    const int rl = q & 63, sl = (rl + (q >> 6)) & 63;
    if (rl == lane) {
      a1 = a1 + V(specRes[0 * specTotal + e], specRes[1 * specTotal + e], specRes[2 * specTotal + e]);
      a2 = a2 + V(specRes[3 * specTotal + e], specRes[4 * specTotal + e], specRes[5 * specTotal + e]);
    }
    if (sl == lane) {
#pragma unroll
      for (int c = 0; c < 6; ++c) cacc[c][lane] += specRes[(6 + c) * specTotal + e];
    }
  }
  double* rp = rowPart + (t - partBase) * 6 * kTile;
  double* cp = colPart + (t - partBase) * 6 * kTile;
  rp[0 * kTile + lane] = a1.x; rp[1 * kTile + lane] = a1.y; rp[2 * kTile + lane] = a1.z;
  rp[3 * kTile + lane] = a2.x; rp[4 * kTile + lane] = a2.y; rp[5 * kTile + lane] = a2.z;
#pragma unroll
  for (int c = 0; c < 6; ++c) cp[c * kTile + lane] = cacc[c][lane];
}

}  // namespace

}  // namespace ddd
