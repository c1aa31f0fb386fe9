// Batched evaluation of listed pairs of a staged 64 x 64 tile, one pair per
// lane, with the results added deterministically to the tile's row sums
// (registers) and column sums (LDS): shared by the hybrid's near-field tiles
// and the mixed-precision tiles.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"
#include "pair_math.hpp"

namespace ddd {

namespace {

// Adds, in increasing position order, the LDS results sres[0..2][e] at the
// set bits e of m.
__device__ __forceinline__ void add_marked(const double (*sres)[kTile], uint64_t m, v3<double>& r) {
  for (; m; m &= m - 1) {
    const int e = __builtin_ctzll(m);
    r = r + V(sres[0][e], sres[1][e], sres[2][e]);
  }
}
__device__ __forceinline__ void add_marked_col(const double (*sres)[kTile], uint64_t m,
                                               double (*cacc)[kTile], int c0, int lane) {
  for (; m; m &= m - 1) {
    const int e = __builtin_ctzll(m);
    cacc[c0][lane] += sres[0][e];
    cacc[c0 + 1][lane] += sres[1][e];
    cacc[c0 + 2][lane] += sres[2][e];
  }
}

// A batch of up to 64 pairs of a staged tile, one per lane (lane e holds
// entry e: row slot e & 63, column slot (row + (e >> 6)) & 63); forces
// F1..F4 from eval.  Each lane then adds, in entry order, the results whose
// row (F1, F2 into a1, a2) or column (F3, F4 into cacc) it owns.  Ownership
// masks are built with LDS atomic ORs; results pass through LDS one force at
// a time.
template <class Eval>
__device__ __forceinline__ void near_batch(unsigned e, int cnt, int lane, Eval eval,
                                           double (*sres)[kTile], double (*cacc)[kTile],
                                           unsigned long long* rowM, unsigned long long* colM,
                                           v3<double>& a1, v3<double>& a2) {
  const int rl = e & 63, sl = (rl + (e >> 6)) & 63;
  rowM[lane] = 0;
  colM[lane] = 0;
  __syncthreads();
  if (lane < cnt) {
    atomicOr(&rowM[rl], 1ull << lane);
    atomicOr(&colM[sl], 1ull << lane);
  }
  __syncthreads();
  const uint64_t mr = rowM[lane], mc = colM[lane];
  v3<double> F1 = V(0.0, 0.0, 0.0), F2 = F1, F3 = F1, F4 = F1;
  if (lane < cnt) eval(rl, sl, F1, F2, F3, F4);
  auto put = [&](const v3<double>& f) {
    sres[0][lane] = f.x;
    sres[1][lane] = f.y;
    sres[2][lane] = f.z;
  };
  put(F1);
  __syncthreads();
  add_marked(sres, mr, a1);
  __syncthreads();
  put(F2);
  __syncthreads();
  add_marked(sres, mr, a2);
  __syncthreads();
  put(F3);
  __syncthreads();
  add_marked_col(sres, mc, cacc, 0, lane);
  __syncthreads();
  put(F4);
  __syncthreads();
  add_marked_col(sres, mc, cacc, 3, lane);
  __syncthreads();
}

// As near_batch, with LDS for two forces of each pair (sres[6][64]): F1 with
// F2, then F3 with F4, in two rounds instead of four.  Every sum receives
// the same values in the same order as in near_batch.
template <class Eval>
__device__ __forceinline__ void near_batch2(unsigned e, int cnt, int lane, Eval eval,
                                            double (*sres)[kTile], double (*cacc)[kTile],
                                            unsigned long long* rowM, unsigned long long* colM,
                                            v3<double>& a1, v3<double>& a2) {
  const int rl = e & 63, sl = (rl + (e >> 6)) & 63;
  rowM[lane] = 0;
  colM[lane] = 0;
  __syncthreads();
  if (lane < cnt) {
    atomicOr(&rowM[rl], 1ull << lane);
    atomicOr(&colM[sl], 1ull << lane);
  }
  __syncthreads();
  const uint64_t mr = rowM[lane], mc = colM[lane];
  v3<double> F1 = V(0.0, 0.0, 0.0), F2 = F1, F3 = F1, F4 = F1;
  if (lane < cnt) eval(rl, sl, F1, F2, F3, F4);
  auto put2 = [&](const v3<double>& f, const v3<double>& g) {
    sres[0][lane] = f.x;
    sres[1][lane] = f.y;
    sres[2][lane] = f.z;
    sres[3][lane] = g.x;
    sres[4][lane] = g.y;
    sres[5][lane] = g.z;
  };
  put2(F1, F2);
  __syncthreads();
  for (uint64_t m = mr; m; m &= m - 1) {
    const int b = __builtin_ctzll(m);
    a1 = a1 + V(sres[0][b], sres[1][b], sres[2][b]);
    a2 = a2 + V(sres[3][b], sres[4][b], sres[5][b]);
  }
  __syncthreads();
  put2(F3, F4);
  __syncthreads();
  for (uint64_t m = mc; m; m &= m - 1) {
    const int b = __builtin_ctzll(m);
#pragma unroll
    for (int c = 0; c < 6; ++c) cacc[c][lane] += sres[c][b];
  }
  __syncthreads();
}

}  // namespace

}  // namespace ddd
