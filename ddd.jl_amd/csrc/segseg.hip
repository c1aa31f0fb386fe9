// FP64 segment-segment force kernels.
//
// Work decomposition: the upper triangle of segment pairs (i < j, the pairs
// calcSegSegForceSer! visits, DislocationForces.jl:794-800) is cut into
// 64 x 64 tiles, grouped 4 x 4 into super-tiles of one workgroup each
// (k_segseg_super).  Lane l of a tile owns row i = 64 I + l and holds its
// node forces in registers; at step k it pairs with column slot
// s = (l + k) mod 64 of the J block, so within one step every lane touches a
// different column and the column forces accumulate in LDS without atomics.
// Each super-tile writes its row and column partial sums; a second kernel
// adds the partials of every segment in a fixed order, so results are
// bitwise reproducible run to run.
//
// Near-parallel pairs (1 - c^2 <= sqrt(eps), the reference's special branch
// :1147-1162) are not evaluated inline, which would stall the whole wavefront
// on one lane: a count pass and a listing pass put them in a per-tile list,
// two packed passes evaluate them one per lane, and the tile adds their
// forces in list order after its own pairs.
#include <algorithm>

#include "kernels.hpp"
#include "pair_special.hpp"
#include "tiles.hpp"

namespace ddd {

namespace {

__global__ void k_prep_segments(double* __restrict__ seg, int64_t ld, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= ld) return;
  if (i >= n) {  // padding: a harmless unit segment, never paired
    for (int f = 0; f < F_COUNT; ++f) seg[f * ld + i] = 0.0;
    seg[F_N2X * ld + i] = 1.0;
    seg[F_TX * ld + i] = 1.0;
    seg[F_TN * ld + i] = 1.0;
    seg[F_EXT * ld + i] = 1.0;
    return;
  }
  v3<double> n1 = V(seg[F_N1X * ld + i], seg[F_N1Y * ld + i], seg[F_N1Z * ld + i]);
  v3<double> n2 = V(seg[F_N2X * ld + i], seg[F_N2Y * ld + i], seg[F_N2Z * ld + i]);
  v3<double> t;
  double tN;
  unit_tangent(n1, n2, t, tN);
  seg[F_TX * ld + i] = t.x;
  seg[F_TY * ld + i] = t.y;
  seg[F_TZ * ld + i] = t.z;
  seg[F_TN * ld + i] = tN;
}

// Near-parallel pairs of the special-pair list, one per lane, forces
// [12][total] (F1, F2 on the row segment, F3, F4 on the column segment):
// pass A of the branch; the pairs that need corrective sub-pairs are listed
// in fix (order of the atomic claims, which does not matter: each pair's
// result has its own slot).
__global__ __launch_bounds__(kTile) void k_special_eval(const double* __restrict__ seg, int64_t ld,
                                                        const unsigned short* __restrict__ specSlot,
                                                        const int* __restrict__ specTile,
                                                        int64_t total, SegK<double> K,
                                                        double* __restrict__ res,
                                                        int64_t* __restrict__ fix,
                                                        unsigned long long* __restrict__ fixCount) {
  const int64_t e = blockIdx.x * (int64_t)kTile + threadIdx.x;
  bool need = false;
  if (e < total) {
    int I, J;
    tile_blocks(specTile[e], I, J);
    const unsigned q = specSlot[e];
    const int rl = q & 63, sl = (rl + (q >> 6)) & 63;
    const Seg<double> P = seg_from(seg, ld, (int64_t)I * kTile + rl);
    const Seg<double> Q = seg_from(seg, ld, (int64_t)J * kTile + sl);
    v3<double> F1, F2, F3, F4;
    need = special_pair_base(K, P, Q, F1, F2, F3, F4);
    const double v[12] = {F1.x, F1.y, F1.z, F2.x, F2.y, F2.z, F3.x, F3.y, F3.z, F4.x, F4.y, F4.z};
#pragma unroll
    for (int c = 0; c < 12; ++c) res[c * total + e] = v[c];
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

__global__ void k_zero_count(unsigned long long* c) { *c = 0; }

// Plain FP64 super-tiles: the S x S (4 x 4) pair tiles of row blocks
// 4P .. 4P + 3 and column blocks 4Q .. 4Q + 3 (P <= Q) in one workgroup of
// four wavefronts.  Wavefront w keeps row block 4P + w for the
// whole super-tile (its row sums in registers, its row segment read into
// registers once per tile) and in round r takes column block
// 4Q + ((w + r) % 4); in every round the four wavefronts hold four
// different column blocks, whose segments and column sums sit in LDS, so no
// two wavefronts ever touch the same sums and a barrier between rounds is
// the only synchronisation across them.  Within a tile, lane l pairs with
// column slot (l + k) % 64 at step k, and the tile's near-parallel pairs,
// evaluated beforehand, are added after its steps in list order.  The
// super-tile reads its 8 blocks of segments once and writes 4 row and 4
// column partial sums (24 KiB) where 16 separate tiles read 32 blocks and
// wrote 96 KiB.  Two workgroups share a CU (64 KiB of LDS each), so the
// two wavefronts of each SIMD come from different workgroups and hide each
// other's LDS latency; one workgroup of eight wavefronts (8 x 8 tiles), its
// pairs of wavefronts kept in step by the round barriers, ran 7 % slower.
// Super-tile u's partials go to slot u - partBase:
// rowPart / colPart [slot][block w][component][lane].
//
// S = 1 is one tile per single-wavefront workgroup (eight per CU): on small
// networks four rounds per workgroup would lengthen the critical path
// fourfold (640 segments: 0.48 ms with one tile per workgroup, 1.38 ms with
// 4 x 4 super-tiles), so those take S = 1.
//
// RANGED: the ranged sequences, done only when all 8 blocks are in the safe
// range and every sum comes out finite; otherwise the super-tile writes
// nothing and returns false (the caller has it redone with the general
// sequences).

// Orders one wavefront's LDS accesses across the steps of a tile: lane l
// adds into column slot (l + k) & 63 at step k, which lane l - 1 updates
// at step k + 1.  A wavefront's LDS operations complete in order, so only
// the compiler has to be kept from moving them across steps.
__device__ __forceinline__ void wave_lds_order() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool RANGED, int S>
__device__ __forceinline__ bool super_eval(int64_t u, const double* __restrict__ seg, int64_t ld,
                                           int n, int nb, const SegK<double>& K,
                                           double* __restrict__ rowPart,
                                           double* __restrict__ colPart,
                                           const long long* __restrict__ specOff,
                                           const unsigned short* __restrict__ specSlot,
                                           const double* __restrict__ specRes, int64_t specTotal,
                                           int64_t partBase, int64_t specBase) {
  __shared__ double si[S][kTileFields][kTile];
  __shared__ double sj[S][kTileFields][kTile];
  __shared__ double cacc[S][6][kTile];
  __shared__ unsigned long long colExtS[S];

  int P, Q;
  tile_blocks(u, P, Q);
  // the wavefront's index, uniform across it (scalar registers, scalar branches)
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kTile);
  const int lane = threadIdx.x % kTile;
  const int I = S * P + w;
  const int Jw = S * Q + w;
  const bool rowValid = I < nb;

  __syncthreads();  // a previous super-tile of this workgroup is done with the arrays
  bool ok = true;
  if (Jw < nb) {  // wavefront w stages column block Jw
    const int64_t j = (int64_t)Jw * kTile + lane;
#pragma unroll
    for (int f = 0; f < kTileFields; ++f) sj[w][f][lane] = seg[f * ld + j];
    const uint64_t m = __ballot(seg[F_EXT * ld + j] != 0.0);
    if (lane == 0) colExtS[w] = m;
    if constexpr (RANGED) {
      const int rf[9] = {F_N1X, F_N1Y, F_N1Z, F_N2X, F_N2Y, F_N2Z, F_BX, F_BY, F_BZ};
#pragma unroll
      for (int f = 0; f < 9; ++f) ok = ok && ranged_value(sj[w][rf[f]][lane]);
    }
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) cacc[w][c][lane] = 0.0;
  bool rowExt = true;
  const int64_t i = (int64_t)I * kTile + lane;
  if (rowValid) {  // and its row block I
#pragma unroll
    for (int f = 0; f < kTileFields; ++f) si[w][f][lane] = seg[f * ld + i];
    rowExt = seg[F_EXT * ld + i] != 0.0;
    if constexpr (RANGED) {
      const int rf[9] = {F_N1X, F_N1Y, F_N1Z, F_N2X, F_N2Y, F_N2Z, F_BX, F_BY, F_BZ};
#pragma unroll
      for (int f = 0; f < 9; ++f) ok = ok && ranged_value(si[w][rf[f]][lane]);
    }
  }
  if constexpr (RANGED) {
    atan_table_fill(threadIdx.x);
    ok = ok && K.aSq >= 0x1p-120 && K.aSq <= 0x1p120;
    if (__syncthreads_or(!ok)) return false;  // also the barrier after staging
  } else {
    __syncthreads();
  }

  const bool rowLive = rowValid && (i < n) && !rowExt;
  v3<double> a1 = V(0.0, 0.0, 0.0), a2 = V(0.0, 0.0, 0.0);
  for (int r = 0; r < S; ++r) {
    const int cb = (w + r) % S;
    const int J = S * Q + cb;
    if (rowValid && J < nb && I <= J) {  // uniform across the wavefront
      const Seg<double> A0 = seg_from_lds(si[w], lane);
      const double (*sjc)[kTile] = sj[cb];
      double (*cc)[kTile] = cacc[cb];
      const uint64_t colExt = colExtS[cb];
      const bool diag = (I == J);
      for (int k = 0; k < kTile; ++k) {
        const int s = (lane + k) % kTile;
        const int64_t j = (int64_t)J * kTile + s;
        const bool live = rowLive && (j < n) && !((colExt >> s) & 1ull) && (!diag || s > lane);
        // the row segment stays in registers across the steps: reading it
        // from LDS at every step exposes that read's latency at the start
        // of the pair (the evaluation took 10 % longer)
        const Seg<double> A = A0;
        const Seg<double> B = seg_from_lds(sjc, s);
        double c;
        const double omcSq = one_minus_c2(A.t, B.t, c);
        if (live && !is_special<0>(true, omcSq, v3<double>{}, B.n1, B.n2, 0.0, 0.0)) {
          v3<double> F1, F2, F3, F4;
          pair_generic<double, true, true, false, RANGED>(K, A, B, c, omcSq, F1, F2, F3, F4);
          a1 = a1 + F1;
          a2 = a2 + F2;
          cc[0][s] += F3.x; cc[1][s] += F3.y; cc[2][s] += F3.z;
          cc[3][s] += F4.x; cc[4][s] += F4.y; cc[5][s] += F4.z;
        }
        wave_lds_order();
      }
      // the tile's near-parallel pairs, in list order.  With several
      // wavefronts per workgroup, 64 entries are loaded at once, one per
      // lane, then taken in turn from the loading lane's registers, so that
      // the other wavefronts do not wait at the round's barrier for one
      // memory latency per entry; a single-wavefront workgroup (S = 1) reads
      // them one at a time, which needs fewer registers, and other
      // workgroups hide the latency.
      const int64_t t = tiles_for(J) + I;
      const long long eb = specOff[t] - specBase, ee = specOff[t + 1] - specBase;
      if constexpr (S == 1) {
        for (long long e = eb; e < ee; ++e) {
          const unsigned qe = specSlot[e];
          const int rl = qe & 63, sl = (rl + (qe >> 6)) & 63;
          if (rl == lane) {
            a1 = a1 + V(specRes[0 * specTotal + e], specRes[1 * specTotal + e],
                        specRes[2 * specTotal + e]);
            a2 = a2 + V(specRes[3 * specTotal + e], specRes[4 * specTotal + e],
                        specRes[5 * specTotal + e]);
          }
          if (sl == lane) {
#pragma unroll
            for (int c = 0; c < 6; ++c) cc[c][lane] += specRes[(6 + c) * specTotal + e];
          }
        }
      }
      for (long long b0 = eb; S > 1 && b0 < ee; b0 += kTile) {
        const int cnt = (int)(ee - b0 < kTile ? ee - b0 : kTile);
        unsigned q = 0;
        double v[12];
#pragma unroll
        for (int c = 0; c < 12; ++c) v[c] = 0.0;
        if (lane < cnt) {
          q = specSlot[b0 + lane];
#pragma unroll
          for (int c = 0; c < 12; ++c) v[c] = specRes[c * specTotal + b0 + lane];
        }
        for (int e = 0; e < cnt; ++e) {
          const unsigned qe = (unsigned)__builtin_amdgcn_readlane((int)q, e);
          const int rl = qe & 63, sl = (rl + (qe >> 6)) & 63;
          double f[12];
#pragma unroll
          for (int c = 0; c < 12; ++c)
            f[c] = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v[c]), e),
                                    __builtin_amdgcn_readlane(__double2loint(v[c]), e));
          if (rl == lane) {
            a1 = a1 + V(f[0], f[1], f[2]);
            a2 = a2 + V(f[3], f[4], f[5]);
          }
          if (sl == lane) {
#pragma unroll
            for (int c = 0; c < 6; ++c) cc[c][lane] += f[6 + c];
          }
        }
      }
    }
    __syncthreads();  // column block cb passes to another wavefront
  }
  if constexpr (RANGED) {
    // inf or NaN anywhere in a row or column sum (a sum overflowing to inf
    // only costs a second evaluation)
    double sum = a1.x + a1.y + a1.z + a2.x + a2.y + a2.z;
#pragma unroll
    for (int c = 0; c < 6; ++c) sum = sum + cacc[w][c][lane];
    if (__syncthreads_or(!(sum - sum == 0.0))) return false;
  }
  double* rp = rowPart + ((u - partBase) * S + w) * 6 * kTile + lane;
  double* cp = colPart + ((u - partBase) * S + w) * 6 * kTile + lane;
  rp[0 * kTile] = a1.x; rp[1 * kTile] = a1.y; rp[2 * kTile] = a1.z;
  rp[3 * kTile] = a2.x; rp[4 * kTile] = a2.y; rp[5 * kTile] = a2.z;
#pragma unroll
  for (int c = 0; c < 6; ++c) cp[c * kTile] = cacc[w][c][lane];
  return true;
}

// Plain FP64 super-tiles with the ranged sequences, one per workgroup
// (u0 + blockIdx.x); those they leave undone are listed in
// redo[0 .. *redoCount) for k_segseg_super_redo.
template <int S>
__global__ __launch_bounds__(S * kTile, 2) void k_segseg_super(
    const double* __restrict__ seg, int64_t ld, int n, int nb, int64_t u0, SegK<double> K,
    double* __restrict__ rowPart, double* __restrict__ colPart,
    const long long* __restrict__ specOff, const unsigned short* __restrict__ specSlot,
    const double* __restrict__ specRes, int64_t specTotal, int64_t partBase, int64_t specBase,
    int* __restrict__ redo, unsigned long long* __restrict__ redoCount) {
  const int64_t u = u0 + blockIdx.x;
  const bool done = super_eval<true, S>(u, seg, ld, n, nb, K, rowPart, colPart, specOff, specSlot,
                                        specRes, specTotal, partBase, specBase);
  if (!done && threadIdx.x == 0) redo[atomicAdd(redoCount, 1ull)] = (int)u;
}

// The super-tiles k_segseg_super left undone, with the general IEEE
// sequences: grid-stride over the list (its length is read on the device).
template <int S>
__global__ __launch_bounds__(S * kTile, 2) void k_segseg_super_redo(
    const double* __restrict__ seg, int64_t ld, int n, int nb, SegK<double> K,
    double* __restrict__ rowPart, double* __restrict__ colPart,
    const long long* __restrict__ specOff, const unsigned short* __restrict__ specSlot,
    const double* __restrict__ specRes, int64_t specTotal, int64_t partBase, int64_t specBase,
    const int* __restrict__ redo, const unsigned long long* __restrict__ redoCount) {
  const int64_t cnt = (int64_t)*redoCount;
  for (int64_t q = blockIdx.x; q < cnt; q += gridDim.x)
    (void)super_eval<false, S>(redo[q], seg, ld, n, nb, K, rowPart, colPart, specOff, specSlot,
                               specRes, specTotal, partBase, specBase);
}

// Sum of a segment's partials of one chunk (super-tiles of S x S tiles,
// super-columns [Q0, Q1), super-tile u at slot u - tiles_for(Q0); S = 1:
// one tile each), in fixed order: its row role in super-tiles (Pb, Q), Q in
// [max(Pb, Q0), Q1), then its column role in super-tiles (P, Pb), P <= Pb,
// when Pb is in the chunk (Pb its block's super-block, w its block's place
// in it).  With one chunk this is every row partial, then every column
// partial of the segment.
template <int S>
__global__ void k_segseg_reduce(const double* __restrict__ rowPart,
                                const double* __restrict__ colPart, int n, int Q0, int Q1,
                                double* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int B = s / kTile, l = s % kTile;
  const int Pb = B / S, w = B % S;
  const int64_t base = tiles_for(Q0);
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int Q = Pb > Q0 ? Pb : Q0; Q < Q1; ++Q) {
    const double* p = rowPart + ((tiles_for(Q) + Pb - base) * S + w) * 6 * kTile + l;
#pragma unroll
    for (int c = 0; c < 6; ++c) acc[c] += p[c * kTile];
  }
  if (Pb >= Q0 && Pb < Q1) {
    const int64_t u0 = tiles_for(Pb) - base;
    for (int P = 0; P <= Pb; ++P) {
      const double* p = colPart + ((u0 + P) * S + w) * 6 * kTile + l;
#pragma unroll
      for (int c = 0; c < 6; ++c) acc[c] += p[c * kTile];
    }
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) out[6 * (int64_t)s + c] += acc[c];
}

// Pass B for the listed pairs: their corrective sub-pairs, added to their
// forces in res.  Grid-stride over the list (its length is on the device).
__global__ __launch_bounds__(kTile) void k_special_fix(const double* __restrict__ seg, int64_t ld,
                                                       const unsigned short* __restrict__ specSlot,
                                                       const int* __restrict__ specTile,
                                                       int64_t total, SegK<double> K,
                                                       double* __restrict__ res,
                                                       const int64_t* __restrict__ fix,
                                                       const unsigned long long* __restrict__ fixCount,
                                                       unsigned* __restrict__ flags) {
  const int64_t cnt = (int64_t)*fixCount;
  for (int64_t q0 = blockIdx.x * (int64_t)kTile + threadIdx.x; q0 < cnt;
       q0 += (int64_t)gridDim.x * kTile) {
    const int64_t e = fix[q0];
    int I, J;
    tile_blocks(specTile[e], I, J);
    const unsigned q = specSlot[e];
    const int rl = q & 63, sl = (rl + (q >> 6)) & 63;
    const Seg<double> P = seg_from(seg, ld, (int64_t)I * kTile + rl);
    const Seg<double> Q = seg_from(seg, ld, (int64_t)J * kTile + sl);
    v3<double> F1 = V(res[0 * total + e], res[1 * total + e], res[2 * total + e]);
    v3<double> F2 = V(res[3 * total + e], res[4 * total + e], res[5 * total + e]);
    v3<double> F3 = V(res[6 * total + e], res[7 * total + e], res[8 * total + e]);
    v3<double> F4 = V(res[9 * total + e], res[10 * total + e], res[11 * total + e]);
    special_pair_fix(K, P, Q, F1, F2, F3, F4, flags);
    const double v[12] = {F1.x, F1.y, F1.z, F2.x, F2.y, F2.z, F3.x, F3.y, F3.z, F4.x, F4.y, F4.z};
#pragma unroll
    for (int c = 0; c < 12; ++c) res[c * total + e] = v[c];
  }
}

// Exclusive scan of n counts (64-bit sums) in three passes: block sums (2,048 per
// workgroup), a one-workgroup scan of the block sums, and the blocks' own
// scans offset by them.  off[n] receives the total.
constexpr int kScanBlock = 2048;

__global__ __launch_bounds__(256) void k_scan_sums(const unsigned* __restrict__ cnt, int n,
                                                   long long* __restrict__ sums) {
  __shared__ long long part[256];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int64_t i0 = (int64_t)b * kScanBlock + 8 * tid;
  long long a = 0;
  for (int q = 0; q < 8; ++q)
    if (i0 + q < n) a += cnt[i0 + q];
  part[tid] = a;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (tid < w) part[tid] += part[tid + w];
    __syncthreads();
  }
  if (tid == 0) sums[b] = part[0];
}

__global__ __launch_bounds__(1024) void k_scan_top(const long long* __restrict__ sums, int nb,
                                                   long long* __restrict__ boff) {
  __shared__ long long part[1024];
  const int tid = threadIdx.x;
  const int per = (nb + 1023) / 1024;
  const int b0 = tid * per, b1 = min(nb, b0 + per);
  long long a = 0;
  for (int b = b0; b < b1; ++b) a += sums[b];
  part[tid] = a;
  __syncthreads();
  if (tid == 0) {
    long long acc = 0;
    for (int k = 0; k < 1024; ++k) {
      const long long v = part[k];
      part[k] = acc;
      acc += v;
    }
    boff[nb] = acc;
  }
  __syncthreads();
  long long acc = part[tid];
  for (int b = b0; b < b1; ++b) {
    boff[b] = acc;
    acc += sums[b];
  }
}

__global__ __launch_bounds__(256) void k_scan_final(const unsigned* __restrict__ cnt, int n,
                                                    const long long* __restrict__ boff, int nb,
                                                    long long* __restrict__ off) {
  __shared__ long long part[256];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int64_t i0 = (int64_t)b * kScanBlock + 8 * tid;
  long long v[8], a = 0;
  for (int q = 0; q < 8; ++q) {
    v[q] = (i0 + q < n) ? (long long)cnt[i0 + q] : 0;
    a += v[q];
  }
  part[tid] = a;
  __syncthreads();
  if (tid == 0) {
    long long acc = boff[b];
    for (int k = 0; k < 256; ++k) {
      const long long x = part[k];
      part[k] = acc;
      acc += x;
    }
  }
  __syncthreads();
  long long acc = part[tid];
  for (int q = 0; q < 8; ++q) {
    if (i0 + q < n) off[i0 + q] = acc;
    acc += v[q];
  }
  if (b == 0 && tid == 0) off[n] = boff[nb];
}

// The list offsets at column-block boundaries: offAt[k] = off[tiles_for(k)].
__global__ void k_block_offsets(const long long* __restrict__ off, int nb,
                                long long* __restrict__ offAt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k <= nb) offAt[k] = off[tiles_for(k)];
}

// Rows-only forces (calcSegSegForceIdx!, :831-877): row r = rows[blockIdx.x]
// against the live segments of column chunk blockIdx.y ([y chunk, (y + 1)
// chunk)), with the row segment as segment 1; one wavefront per (row,
// chunk), its lanes striding over the chunk, reduced across lanes in a fixed
// butterfly order into part[row][chunk][6] (or, with one chunk and part
// NULL, added into out[6 row .. 6 row + 5]).  Splitting the row over chunks
// keeps the latency of a short idx list (a topology update re-evaluates a
// few rows) at a few pairs per lane instead of n / 64.
__global__ __launch_bounds__(kTile, 2) void k_segseg_rows(const double* __restrict__ seg, int64_t ld,
                                                       int n, const int64_t* __restrict__ rows,
                                                       int chunk, SegK<double> K,
                                                       double* __restrict__ part,
                                                       double* __restrict__ out,
                                                       unsigned* __restrict__ flags) {
  const int64_t w = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t r = rows[w];
  const int64_t j0 = (int64_t)blockIdx.y * chunk;
  const int64_t j1 = j0 + chunk < n ? j0 + chunk : n;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  if (r >= 0 && r < n && seg[F_EXT * ld + r] == 0.0) {
    const Seg<double> A = seg_from(seg, ld, r);
    for (int64_t j = j0 + lane; j < j1; j += kTile) {
      if (j == r || seg[F_EXT * ld + j] != 0.0) continue;
      const Seg<double> B = seg_from(seg, ld, j);
      double c;
      const double omcSq = one_minus_c2(A.t, B.t, c);
      v3<double> F1, F2, F3, F4;
      if (omcSq > DDD_SQRT_EPS) {
        pair_generic<double, true, false>(K, A, B, c, omcSq, F1, F2, F3, F4);
      } else {
        special_pair_dev(K, A, B, F1, F2, F3, F4, flags);
      }
      acc[0] += F1.x; acc[1] += F1.y; acc[2] += F1.z;
      acc[3] += F2.x; acc[4] += F2.y; acc[5] += F2.z;
    }
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    double v = acc[c];
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kTile);
    acc[c] = v;
  }
  if (lane == 0) {
    if (part) {
      double* p = part + (w * gridDim.y + blockIdx.y) * 6;
#pragma unroll
      for (int c = 0; c < 6; ++c) p[c] = acc[c];
    } else {
#pragma unroll
      for (int c = 0; c < 6; ++c) out[6 * w + c] += acc[c];
    }
  }
}

// out[6 w + c] += the row's chunk partials, in chunk order.
__global__ void k_segseg_rows_sum(const double* __restrict__ part, int nRows, int nChunk,
                                  double* __restrict__ out) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= (int64_t)nRows * 6) return;
  const int64_t w = e / 6, c = e % 6;
  double a = 0.0;
  for (int q = 0; q < nChunk; ++q) a += part[(w * nChunk + q) * 6 + c];
  out[e] += a;
}

// The tile kernel's ranged sequences on given arguments, for tests that
// compare them bit for bit with the host's general ones: y[i] =
// log_ranged(x[i]) (which 0), atan_ranged(x[i]) (1), sqrt_ranged(x[i]) (2),
// div_ranged(x[2i], x[2i+1]) (3).  Grid-stride over i.
__global__ __launch_bounds__(kTile) void k_ranged_selftest(int which, const double* __restrict__ x,
                                                           int64_t n, double* __restrict__ y) {
  atan_table_fill(threadIdx.x);
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)kTile + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kTile) {
    double r;
    if (which == 0) r = log_ranged(x[i]);
    else if (which == 1) r = atan_ranged(x[i]);
    else if (which == 2) r = sqrt_ranged(x[i]);
    else r = div_ranged(x[2 * i], x[2 * i + 1]);
    y[i] = r;
  }
}

}  // namespace

void launch_prep_segments(double* seg, int64_t ld, int64_t n, hipStream_t st) {
  const int bs = 256;
  hipLaunchKernelGGL(k_prep_segments, dim3((unsigned)((ld + bs - 1) / bs)), dim3(bs), 0, st, seg,
                     ld, n);
}

template <int LIST>
static void special_pass_plain(const double* seg, int64_t ld, int n, unsigned* cnt,
                               const SpecialList* sp, int64_t t0, int64_t t1, hipStream_t st) {
  const int64_t maxGrid = 1ll << 30;
  for (int64_t a = t0; a < t1; a += maxGrid) {
    const int64_t count = (t1 - a < maxGrid) ? (t1 - a) : maxGrid;
    hipLaunchKernelGGL((k_special_pass<0, LIST>), dim3((unsigned)count), dim3(kTile), 0, st, seg, ld, n,
                       a, 0.0, 0.0, cnt, sp ? sp->off : nullptr, sp ? sp->base : 0,
                       sp ? sp->slot : nullptr, sp ? sp->tile : nullptr);
  }
}

void launch_special_count(const double* seg, int64_t ld, int n, const MixedRule* mixed,
                          unsigned* cnt, hipStream_t st) {
  const int64_t T = tiles_for((n + kTile - 1) / kTile);
  if (mixed) launch_special_pass_mixed(seg, ld, n, *mixed, cnt, nullptr, 0, T, st);
  else special_pass_plain<0>(seg, ld, n, cnt, nullptr, 0, T, st);
}

int64_t scan_work_ints(int n) { return 2 * ((int64_t)(n + kScanBlock - 1) / kScanBlock + 1); }

void launch_exclusive_scan(const unsigned* cnt, int n, long long* off, long long* work,
                           hipStream_t st) {
  const int nb = (n + kScanBlock - 1) / kScanBlock;
  long long* sums = work;
  long long* boff = work + nb + 1;
  if (nb > 0) hipLaunchKernelGGL(k_scan_sums, dim3(nb), dim3(256), 0, st, cnt, n, sums);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, sums, nb, boff);
  hipLaunchKernelGGL(k_scan_final, dim3(nb > 0 ? nb : 1), dim3(256), 0, st, cnt, n, boff, nb, off);
}

void launch_block_offsets(const long long* off, int nb, long long* offAt, hipStream_t st) {
  hipLaunchKernelGGL(k_block_offsets, dim3((nb + 1 + 255) / 256), dim3(256), 0, st, off, nb, offAt);
}

void launch_special_list(const double* seg, int64_t ld, int n, const MixedRule* mixed,
                         const SpecialList& sp, const SegK<double>& K, const TileRange& r,
                         hipStream_t st) {
  if (sp.total <= 0) return;
  if (mixed) launch_special_pass_mixed(seg, ld, n, *mixed, nullptr, &sp, r.t0(), r.t1(), st);
  else special_pass_plain<1>(seg, ld, n, nullptr, &sp, r.t0(), r.t1(), st);
  hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(1), 0, st, sp.fixCount);
  hipLaunchKernelGGL(k_special_eval, dim3((unsigned)((sp.total + kTile - 1) / kTile)), dim3(kTile), 0,
                     st, seg, ld, sp.slot, sp.tile, sp.total, K, sp.res, sp.fix, sp.fixCount);
  const int64_t blocks = std::min<int64_t>((sp.total + kTile - 1) / kTile, 2048);
  hipLaunchKernelGGL(k_special_fix, dim3((unsigned)blocks), dim3(kTile), 0, st, seg, ld, sp.slot,
                     sp.tile, sp.total, K, sp.res, sp.fix, sp.fixCount, sp.flags);
}

void launch_segseg_tiles(const double* seg, int64_t ld, int n, const SegK<double>& K,
                         double* rowPart, double* colPart, const SpecialList& sp,
                         const TileRange& r, hipStream_t st, int* redo,
                         unsigned long long* redoCount, const MixedRule* mixed,
                         unsigned short* stage, unsigned long long* counters) {
  if (mixed) {
    launch_segseg_tiles_mixed(seg, ld, n, K, rowPart, colPart, sp, r, st, *mixed, stage, counters);
    return;
  }
  hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(1), 0, st, redoCount);
  const int nb = (n + kTile - 1) / kTile;
  const int64_t u0 = r.u0(), u1 = r.u1();
  // with an empty list each redo workgroup reads the count and ends
  const int64_t blocks = std::min<int64_t>(u1 - u0, 2048);
  if (r.S == kSuper) {
    hipLaunchKernelGGL(k_segseg_super<kSuper>, dim3((unsigned)(u1 - u0)), dim3(kSuper * kTile), 0,
                       st, seg, ld, n, nb, u0, K, rowPart, colPart, sp.off, sp.slot, sp.res,
                       sp.total, u0, sp.base, redo, redoCount);
    hipLaunchKernelGGL(k_segseg_super_redo<kSuper>, dim3((unsigned)blocks), dim3(kSuper * kTile), 0,
                       st, seg, ld, n, nb, K, rowPart, colPart, sp.off, sp.slot, sp.res, sp.total,
                       u0, sp.base, redo, redoCount);
  } else {
    hipLaunchKernelGGL(k_segseg_super<1>, dim3((unsigned)(u1 - u0)), dim3(kTile), 0, st, seg, ld,
                       n, nb, u0, K, rowPart, colPart, sp.off, sp.slot, sp.res, sp.total, u0,
                       sp.base, redo, redoCount);
    hipLaunchKernelGGL(k_segseg_super_redo<1>, dim3((unsigned)blocks), dim3(kTile), 0, st, seg, ld,
                       n, nb, K, rowPart, colPart, sp.off, sp.slot, sp.res, sp.total, u0, sp.base,
                       redo, redoCount);
  }
}

void launch_segseg_reduce(const double* rowPart, const double* colPart, int n, const TileRange& r,
                          double* out, hipStream_t st) {
  // segments of row blocks below Q1 S receive partials from this chunk
  const int64_t m = std::min<int64_t>(n, (int64_t)r.Q1() * r.S * kTile);
  if (m <= 0) return;
  const int bs = 256;
  if (r.S == kSuper)
    hipLaunchKernelGGL(k_segseg_reduce<kSuper>, dim3((unsigned)((m + bs - 1) / bs)), dim3(bs), 0, st,
                       rowPart, colPart, (int)m, r.Q0(), r.Q1(), out);
  else
    hipLaunchKernelGGL(k_segseg_reduce<1>, dim3((unsigned)((m + bs - 1) / bs)), dim3(bs), 0, st,
                       rowPart, colPart, (int)m, r.Q0(), r.Q1(), out);
}

void launch_ranged_selftest(int which, const double* x, int64_t n, double* y, hipStream_t st) {
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>((n + kTile - 1) / kTile, 4096);
  hipLaunchKernelGGL(k_ranged_selftest, dim3((unsigned)blocks), dim3(kTile), 0, st, which, x, n, y);
}

int segseg_rows_chunks(int n, int nRows) {
  if (nRows <= 0 || n <= 0) return 1;
  // about 4,096 wavefronts in all, but at least 256 columns (4 per lane) each
  const int64_t want = (4096 + nRows - 1) / nRows;
  const int64_t most = ((int64_t)n + 255) / 256;
  return (int)std::max<int64_t>(1, std::min(want, most));
}

void launch_segseg_rows(const double* seg, int64_t ld, int n, const int64_t* rows, int nRows,
                        const SegK<double>& K, double* out, double* part, int nChunk,
                        unsigned* flags, hipStream_t st) {
  if (nRows <= 0) return;
  const int chunk = (int)((((int64_t)n + nChunk - 1) / nChunk + kTile - 1) / kTile * kTile);
  if (nChunk == 1) {
    hipLaunchKernelGGL(k_segseg_rows, dim3((unsigned)nRows, 1u), dim3(kTile), 0, st, seg, ld, n, rows,
                       chunk, K, nullptr, out, flags);
    return;
  }
  hipLaunchKernelGGL(k_segseg_rows, dim3((unsigned)nRows, (unsigned)nChunk), dim3(kTile), 0, st, seg,
                     ld, n, rows, chunk, K, part, out, flags);
  const int64_t m = (int64_t)nRows * 6;
  hipLaunchKernelGGL(k_segseg_rows_sum, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, part,
                     nRows, nChunk, out);
}

}  // namespace ddd
