// FP64 segment-segment force kernels.
//
// Work decomposition: the upper triangle of segment pairs (i < j, the pairs
// calcSegSegForceSer! visits, DislocationForces.jl:794-800) is cut into
// 64 x 64 tiles, one wavefront per tile.  Lane l owns row i = 64 I + l and
// holds its node forces in registers; at step k it pairs with column slot
// s = (l + k) mod 64 of the J block, so within one step every lane touches a
// different column and the column forces accumulate in LDS without atomics.
// Each tile writes its row and column partial sums; a second kernel adds the
// partials of every segment in a fixed order, so results are bitwise
// reproducible run to run.
//
// Near-parallel pairs (1 - c^2 <= sqrt(eps), the reference's special branch
// :1147-1162) are not evaluated inline, which would stall the whole wavefront
// on one lane.  They are queued in LDS and evaluated 64 at a time, one per
// lane, then added to their row and column in queue order.
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

constexpr int kQueue = 2 * kTile;

// Tile t of the upper triangle, enumerated column-block major:
// t = J (J + 1) / 2 + I with 0 <= I <= J.
__device__ __forceinline__ void tile_blocks(int64_t t, int& I, int& J) {
  J = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((int64_t)J * (J + 1) / 2 > t) --J;
  while ((int64_t)(J + 1) * (J + 2) / 2 <= t) ++J;
  I = (int)(t - (int64_t)J * (J + 1) / 2);
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
  const double dx = 0.5 * (n1B.x + n2B.x) - midA.x, dy = 0.5 * (n1B.y + n2B.y) - midA.y,
               dz = 0.5 * (n1B.z + n2B.z) - midA.z;
  const bool far = (dx * dx + dy * dy + dz * dz > farDist2) && (omcSq > sinMin2);
  return !far;
}

// Near-parallel pairs per tile, one wavefront per tile in the tile kernel's
// slot order (k, lane): LIST = 0 counts them (cnt[t]); LIST = 1 writes them
// to the tile's range off[t] .. off[t + 1] of the pair list.
template <int MODE, int LIST>
__global__ __launch_bounds__(kTile) void k_special_pass(const double* __restrict__ seg, int64_t ld,
                                                        int n, int64_t tile0, double farDist2,
                                                        double sinMin2, unsigned* __restrict__ cnt,
                                                        const int* __restrict__ off,
                                                        unsigned short* __restrict__ slot,
                                                        int* __restrict__ tileOf) {
  const int64_t t = tile0 + blockIdx.x;
  int I, J;
  tile_blocks(t, I, J);
  const int lane = threadIdx.x;
  __shared__ double sj[10][kTile];  // n1, n2, t, ext
  const int64_t i = (int64_t)I * kTile + lane;
  const int64_t jl = (int64_t)J * kTile + lane;
  const int fl[10] = {F_N1X, F_N1Y, F_N1Z, F_N2X, F_N2Y, F_N2Z, F_TX, F_TY, F_TZ, F_EXT};
#pragma unroll
  for (int f = 0; f < 10; ++f) sj[f][lane] = seg[fl[f] * ld + jl];
  __syncthreads();
  const v3<double> tA = V(seg[F_TX * ld + i], seg[F_TY * ld + i], seg[F_TZ * ld + i]);
  const v3<double> midA = V(0.5 * (seg[F_N1X * ld + i] + seg[F_N2X * ld + i]),
                            0.5 * (seg[F_N1Y * ld + i] + seg[F_N2Y * ld + i]),
                            0.5 * (seg[F_N1Z * ld + i] + seg[F_N2Z * ld + i]));
  const bool rowLive = (i < n) && seg[F_EXT * ld + i] == 0.0;
  const bool diag = (I == J);
  const uint64_t ltMask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int c = LIST ? off[t] : 0;
  const int cEnd = LIST ? off[t + 1] : 0;
  for (int k = 0; k < kTile; ++k) {
    const int s = (lane + k) & 63;
    const int64_t j = (int64_t)J * kTile + s;
    const bool live = rowLive && (j < n) && sj[9][s] == 0.0 && (!diag || s > lane);
    double cc;
    const double omcSq = one_minus_c2(tA, V(sj[6][s], sj[7][s], sj[8][s]), cc);
    const bool sp = is_special<MODE>(live, omcSq, midA, V(sj[0][s], sj[1][s], sj[2][s]),
                                     V(sj[3][s], sj[4][s], sj[5][s]), farDist2, sinMin2);
    const uint64_t m = __ballot(sp);
    if (LIST && sp) {
      const int pos = c + __popcll(m & ltMask);
      if (pos < cEnd) {  // always true: both passes classify identically
        slot[pos] = (unsigned short)((k << 6) | lane);
        tileOf[pos] = (int)t;
      }
    }
    c += __popcll(m);
  }
  if (!LIST && lane == 0) cnt[t] = (unsigned)c;
}

// Pair-tile kernel.  MODE 0 (plain FP64): every generic pair runs inline in
// FP64.  MODE 1 (mixed precision): well-separated pairs (midpoint distance >
// farDist and |t1 x t2| > sinMin) run inline in FP32 by quadrature
// (far_side); the other generic pairs are queued and evaluated in FP64 64 at
// a time, so the wavefront never waits on a minority path.  In both modes
// near-parallel pairs (the reference's special branch, :1147-1162, several
// times the cost of a generic pair) are not evaluated here: they were listed
// per tile and evaluated beforehand (k_special_pass, k_special_eval), and
// the tile adds its list's forces, in list order, after its own pairs.
// Keeping that branch out of this kernel lets it run two wavefronts per
// SIMD.  The number of FP64 pairs is added to *nExact (mixed mode).
template <int MODE>
__global__ __launch_bounds__(kTile, 2) void k_segseg_tiles(const double* __restrict__ seg,
                                                           int64_t ld, int n, int64_t tile0,
                                                           SegK<double> K, SigK<float> KF,
                                                           double farDist2, double sinMin2,
                                                           double* __restrict__ rowPart,
                                                           double* __restrict__ colPart,
                                                           const int* __restrict__ specOff,
                                                           const unsigned short* __restrict__ specSlot,
                                                           const double* __restrict__ specRes,
                                                           int64_t specTotal,
                                                           unsigned long long* __restrict__ nExact) {
  const int64_t t = tile0 + blockIdx.x;
  int I, J;
  tile_blocks(t, I, J);
  const int lane = threadIdx.x;

  __shared__ double si[F_COUNT][kTile];
  __shared__ double sj[F_COUNT][kTile];
  __shared__ double cacc[6][kTile];
  __shared__ double sres[12][kTile];
  __shared__ unsigned short queue[kQueue];

  const int64_t i = (int64_t)I * kTile + lane;
  const int64_t jl = (int64_t)J * kTile + lane;
#pragma unroll
  for (int f = 0; f < F_COUNT; ++f) {
    si[f][lane] = seg[f * ld + i];
    sj[f][lane] = seg[f * ld + jl];
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) cacc[c][lane] = 0.0;
  __syncthreads();

  const Seg<double> A = seg_from_lds(si, lane);
  const bool rowLive = (i < n) && si[F_EXT][lane] == 0.0;
  const bool diag = (I == J);
  v3<double> a1 = V(0.0, 0.0, 0.0), a2 = V(0.0, 0.0, 0.0);
  int nq = 0;        // wave-uniform queue length (mixed mode)

  unsigned exact = 0;
  const uint64_t ltMask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const v3<double> midA = V(0.5 * (A.n1.x + A.n2.x), 0.5 * (A.n1.y + A.n2.y), 0.5 * (A.n1.z + A.n2.z));

  // Mixed mode: evaluate the first cnt queued generic pairs in FP64 and add
  // them to rows and columns in queue order.
  auto flush = [&](int cnt) {
    if (lane < cnt) {
      const unsigned e = queue[lane];
      const int rl = e & 63, sl = (rl + (e >> 6)) & 63;
      const Seg<double> P = seg_from_lds(si, rl), Q = seg_from_lds(sj, sl);
      v3<double> F1, F2, F3, F4;
      double c;
      const double omc = one_minus_c2(P.t, Q.t, c);
      pair_generic<double, true, true>(K, P, Q, c, omc, F1, F2, F3, F4);
      sres[0][lane] = F1.x; sres[1][lane] = F1.y; sres[2][lane] = F1.z;
      sres[3][lane] = F2.x; sres[4][lane] = F2.y; sres[5][lane] = F2.z;
      sres[6][lane] = F3.x; sres[7][lane] = F3.y; sres[8][lane] = F3.z;
      sres[9][lane] = F4.x; sres[10][lane] = F4.y; sres[11][lane] = F4.z;
    }
    __syncthreads();
    for (int e = 0; e < cnt; ++e) {  // queue order: deterministic
      const unsigned q = queue[e];
      const int rl = q & 63, sl = (rl + (q >> 6)) & 63;
      if (rl == lane) {
        a1 = a1 + V(sres[0][e], sres[1][e], sres[2][e]);
        a2 = a2 + V(sres[3][e], sres[4][e], sres[5][e]);
      }
      if (sl == lane) {
#pragma unroll
        for (int c = 0; c < 6; ++c) cacc[c][lane] += sres[6 + c][e];
      }
    }
    __syncthreads();
  };

  for (int k = 0; k < kTile; ++k) {
    const int s = (lane + k) & 63;
    const int64_t j = (int64_t)J * kTile + s;
    const bool live = rowLive && (j < n) && sj[F_EXT][s] == 0.0 && (!diag || s > lane);
    const Seg<double> B = seg_from_lds(sj, s);
    double c;
    const double omcSq = one_minus_c2(A.t, B.t, c);
    const bool special = is_special<MODE>(live, omcSq, midA, B.n1, B.n2, farDist2, sinMin2);
    bool queued = false, inlineFP64 = false;
    if (MODE == 0) {
      inlineFP64 = live && !special;
    } else {
      const double dx = 0.5 * (B.n1.x + B.n2.x) - midA.x, dy = 0.5 * (B.n1.y + B.n2.y) - midA.y,
                   dz = 0.5 * (B.n1.z + B.n2.z) - midA.z;
      const bool far = (dx * dx + dy * dy + dz * dz > farDist2) && (omcSq > sinMin2);
      queued = live && !far && !special;
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
    }
    if (MODE == 1) {
      const uint64_t m = __ballot(queued);
      if (m) {
        if (queued) queue[nq + __popcll(m & ltMask)] = (unsigned short)((k << 6) | lane);
        nq += __popcll(m);
        exact += __popcll(m);
      }
    }
    if (MODE == 0 && inlineFP64) {
      v3<double> F1, F2, F3, F4;
      pair_generic<double, true, true>(K, A, B, c, omcSq, F1, F2, F3, F4);
      a1 = a1 + F1;
      a2 = a2 + F2;
      cacc[0][s] += F3.x; cacc[1][s] += F3.y; cacc[2][s] += F3.z;
      cacc[3][s] += F4.x; cacc[4][s] += F4.y; cacc[5][s] += F4.z;
    }
    __syncthreads();
    if (MODE == 1 && nq >= kTile) {
      flush(kTile);
      const int rest = nq - kTile;
      unsigned short keep = (lane < rest) ? queue[kTile + lane] : 0;
      __syncthreads();
      if (lane < rest) queue[lane] = keep;
      __syncthreads();
      nq = rest;
    }
  }
  if (MODE == 1 && nq > 0) flush(nq);
  if (MODE == 1 && lane == 0 && exact) atomicAdd(nExact, (unsigned long long)exact);
  // this tile's near-parallel pairs, evaluated beforehand, in list order
  for (int e = specOff[t], e1 = specOff[t + 1]; e < e1; ++e) {
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

  double* rp = rowPart + t * 6 * kTile;
  double* cp = colPart + t * 6 * kTile;
  rp[0 * kTile + lane] = a1.x; rp[1 * kTile + lane] = a1.y; rp[2 * kTile + lane] = a1.z;
  rp[3 * kTile + lane] = a2.x; rp[4 * kTile + lane] = a2.y; rp[5 * kTile + lane] = a2.z;
#pragma unroll
  for (int c = 0; c < 6; ++c) cp[c * kTile + lane] = cacc[c][lane];
}

// Near-parallel pairs of the special-pair list, one per lane, forces
// [12][total] (F1, F2 on the row segment, F3, F4 on the column segment).
__global__ __launch_bounds__(kTile) void k_special_eval(const double* __restrict__ seg, int64_t ld,
                                                        const unsigned short* __restrict__ specSlot,
                                                        const int* __restrict__ specTile,
                                                        int64_t total, SegK<double> K,
                                                        double* __restrict__ res,
                                                        unsigned* __restrict__ flags) {
  const int64_t e = blockIdx.x * (int64_t)kTile + threadIdx.x;
  if (e >= total) return;
  int I, J;
  tile_blocks(specTile[e], I, J);
  const unsigned q = specSlot[e];
  const int rl = q & 63, sl = (rl + (q >> 6)) & 63;
  const Seg<double> P = seg_from(seg, ld, (int64_t)I * kTile + rl);
  const Seg<double> Q = seg_from(seg, ld, (int64_t)J * kTile + sl);
  v3<double> F1, F2, F3, F4;
  special_pair_dev(K, P, Q, F1, F2, F3, F4, flags);
  const double v[12] = {F1.x, F1.y, F1.z, F2.x, F2.y, F2.z, F3.x, F3.y, F3.z, F4.x, F4.y, F4.z};
#pragma unroll
  for (int c = 0; c < 12; ++c) res[c * total + e] = v[c];
}

// Exclusive scan of n counts in three passes: block sums (2,048 per
// workgroup), a one-workgroup scan of the block sums, and the blocks' own
// scans offset by them.  off[n] receives the total.
constexpr int kScanBlock = 2048;

__global__ __launch_bounds__(256) void k_scan_sums(const unsigned* __restrict__ cnt, int n,
                                                   int* __restrict__ sums) {
  __shared__ int part[256];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int64_t i0 = (int64_t)b * kScanBlock + 8 * tid;
  int a = 0;
  for (int q = 0; q < 8; ++q)
    if (i0 + q < n) a += (int)cnt[i0 + q];
  part[tid] = a;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (tid < w) part[tid] += part[tid + w];
    __syncthreads();
  }
  if (tid == 0) sums[b] = part[0];
}

__global__ __launch_bounds__(1024) void k_scan_top(const int* __restrict__ sums, int nb,
                                                   int* __restrict__ boff) {
  __shared__ int part[1024];
  const int tid = threadIdx.x;
  const int per = (nb + 1023) / 1024;
  const int b0 = tid * per, b1 = min(nb, b0 + per);
  int a = 0;
  for (int b = b0; b < b1; ++b) a += sums[b];
  part[tid] = a;
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int k = 0; k < 1024; ++k) {
      const int v = part[k];
      part[k] = acc;
      acc += v;
    }
    boff[nb] = acc;
  }
  __syncthreads();
  int acc = part[tid];
  for (int b = b0; b < b1; ++b) {
    boff[b] = acc;
    acc += sums[b];
  }
}

__global__ __launch_bounds__(256) void k_scan_final(const unsigned* __restrict__ cnt, int n,
                                                    const int* __restrict__ boff, int nb,
                                                    int* __restrict__ off) {
  __shared__ int part[256];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int64_t i0 = (int64_t)b * kScanBlock + 8 * tid;
  int v[8], a = 0;
  for (int q = 0; q < 8; ++q) {
    v[q] = (i0 + q < n) ? (int)cnt[i0 + q] : 0;
    a += v[q];
  }
  part[tid] = a;
  __syncthreads();
  if (tid == 0) {
    int acc = boff[b];
    for (int k = 0; k < 256; ++k) {
      const int x = part[k];
      part[k] = acc;
      acc += x;
    }
  }
  __syncthreads();
  int acc = part[tid];
  for (int q = 0; q < 8; ++q) {
    if (i0 + q < n) off[i0 + q] = acc;
    acc += v[q];
  }
  if (b == 0 && tid == 0) off[n] = boff[nb];
}

// Sum of a segment's partials: row role in tiles (B, J), J >= B, then column
// role in tiles (I, B), I <= B.
__global__ void k_segseg_reduce(const double* __restrict__ rowPart,
                                const double* __restrict__ colPart, int n, int nb,
                                double* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int B = s / kTile, l = s % kTile;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int J = B; J < nb; ++J) {
    const double* p = rowPart + (tiles_for(J) + B) * 6 * kTile + l;
#pragma unroll
    for (int c = 0; c < 6; ++c) acc[c] += p[c * kTile];
  }
  const int64_t t0 = tiles_for(B);
  for (int I = 0; I <= B; ++I) {
    const double* p = colPart + (t0 + I) * 6 * kTile + l;
#pragma unroll
    for (int c = 0; c < 6; ++c) acc[c] += p[c * kTile];
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) out[6 * (int64_t)s + c] += acc[c];
}

// Rows-only forces (calcSegSegForceIdx!, :831-877): row r = rows[w] against
// every other live segment, with the row segment as segment 1; one
// wavefront per row, reduced across lanes in a fixed butterfly order.
__global__ __launch_bounds__(kTile) void k_segseg_rows(const double* __restrict__ seg, int64_t ld,
                                                       int n, const int64_t* __restrict__ rows,
                                                       SegK<double> K, double* __restrict__ out,
                                                       unsigned* __restrict__ flags) {
  const int w = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t r = rows[w];
  double acc[6] = {0, 0, 0, 0, 0, 0};
  if (r >= 0 && r < n && seg[F_EXT * ld + r] == 0.0) {
    const Seg<double> A = seg_from(seg, ld, r);
    for (int64_t j = lane; j < n; j += kTile) {
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
#pragma unroll
    for (int c = 0; c < 6; ++c) out[6 * (int64_t)w + c] += acc[c];
  }
}

}  // namespace

void launch_prep_segments(double* seg, int64_t ld, int64_t n, hipStream_t st) {
  const int bs = 256;
  hipLaunchKernelGGL(k_prep_segments, dim3((unsigned)((ld + bs - 1) / bs)), dim3(bs), 0, st, seg,
                     ld, n);
}

template <int LIST>
static void special_pass(const double* seg, int64_t ld, int n, const MixedRule* mixed, unsigned* cnt,
                         const SpecialList* sp, hipStream_t st) {
  const int64_t nb = (n + kTile - 1) / kTile;
  const int64_t T = tiles_for(nb);
  const int64_t maxGrid = 1ll << 30;
  const double fd2 = mixed ? mixed->farDist * mixed->farDist : 0.0;
  const double sm2 = mixed ? mixed->sinMin * mixed->sinMin : 0.0;
  const int* off = sp ? sp->off : nullptr;
  unsigned short* slot = sp ? sp->slot : nullptr;
  int* tileOf = sp ? sp->tile : nullptr;
  for (int64_t t0 = 0; t0 < T; t0 += maxGrid) {
    const int64_t count = (T - t0 < maxGrid) ? (T - t0) : maxGrid;
    if (mixed)
      hipLaunchKernelGGL((k_special_pass<1, LIST>), dim3((unsigned)count), dim3(kTile), 0, st, seg, ld,
                         n, t0, fd2, sm2, cnt, off, slot, tileOf);
    else
      hipLaunchKernelGGL((k_special_pass<0, LIST>), dim3((unsigned)count), dim3(kTile), 0, st, seg, ld,
                         n, t0, fd2, sm2, cnt, off, slot, tileOf);
  }
}

void launch_special_count(const double* seg, int64_t ld, int n, const MixedRule* mixed,
                          unsigned* cnt, hipStream_t st) {
  special_pass<0>(seg, ld, n, mixed, cnt, nullptr, st);
}

int64_t scan_work_ints(int n) { return 2 * ((int64_t)(n + kScanBlock - 1) / kScanBlock + 1); }

void launch_exclusive_scan(const unsigned* cnt, int n, int* off, int* work, hipStream_t st) {
  const int nb = (n + kScanBlock - 1) / kScanBlock;
  int* sums = work;
  int* boff = work + nb + 1;
  if (nb > 0) hipLaunchKernelGGL(k_scan_sums, dim3(nb), dim3(256), 0, st, cnt, n, sums);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, sums, nb, boff);
  hipLaunchKernelGGL(k_scan_final, dim3(nb > 0 ? nb : 1), dim3(256), 0, st, cnt, n, boff, nb, off);
}

void launch_special_list(const double* seg, int64_t ld, int n, const MixedRule* mixed,
                         const SpecialList& sp, const SegK<double>& K, hipStream_t st) {
  if (sp.total <= 0) return;
  special_pass<1>(seg, ld, n, mixed, nullptr, &sp, st);
  hipLaunchKernelGGL(k_special_eval, dim3((unsigned)((sp.total + kTile - 1) / kTile)), dim3(kTile), 0,
                     st, seg, ld, sp.slot, sp.tile, sp.total, K, sp.res, sp.flags);
}

void launch_segseg_tiles(const double* seg, int64_t ld, int n, const SegK<double>& K,
                         double* rowPart, double* colPart, const SpecialList& sp, hipStream_t st,
                         const MixedRule* mixed, unsigned long long* nExact) {
  const int64_t nb = (n + kTile - 1) / kTile;
  const int64_t T = tiles_for(nb);
  const int64_t maxGrid = 1ll << 30;
  const SigK<float> KF = mixed ? mixed->K : SigK<float>{};
  const double fd2 = mixed ? mixed->farDist * mixed->farDist : 0.0;
  const double sm2 = mixed ? mixed->sinMin * mixed->sinMin : 0.0;
  for (int64_t t0 = 0; t0 < T; t0 += maxGrid) {
    const int64_t cnt = (T - t0 < maxGrid) ? (T - t0) : maxGrid;
    if (mixed)
      hipLaunchKernelGGL(k_segseg_tiles<1>, dim3((unsigned)cnt), dim3(kTile), 0, st, seg, ld, n, t0,
                         K, KF, fd2, sm2, rowPart, colPart, sp.off, sp.slot, sp.res, sp.total, nExact);
    else
      hipLaunchKernelGGL(k_segseg_tiles<0>, dim3((unsigned)cnt), dim3(kTile), 0, st, seg, ld, n, t0,
                         K, KF, fd2, sm2, rowPart, colPart, sp.off, sp.slot, sp.res, sp.total, nExact);
  }
}

void launch_segseg_reduce(const double* rowPart, const double* colPart, int n, double* out,
                          hipStream_t st) {
  const int nb = (n + kTile - 1) / kTile;
  const int bs = 256;
  hipLaunchKernelGGL(k_segseg_reduce, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, st,
                     rowPart, colPart, n, nb, out);
}

void launch_segseg_rows(const double* seg, int64_t ld, int n, const int64_t* rows, int nRows,
                        const SegK<double>& K, double* out, unsigned* flags, hipStream_t st) {
  if (nRows <= 0) return;
  hipLaunchKernelGGL(k_segseg_rows, dim3((unsigned)nRows), dim3(kTile), 0, st, seg, ld, n, rows, K,
                     out, flags);
}

}  // namespace ddd
