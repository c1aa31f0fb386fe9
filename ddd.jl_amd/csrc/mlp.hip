// The far-field network (no reference counterpart): a 4-layer MLP mapping a
// cell offset o to the 60 x 90 operator W(o) of the hybrid far field,
//   phi(o) (128) -> 512 -> 512 -> 512 -> 5408 (5400 used), SiLU activations,
// W(o)[q][k] = env(o, q, k) * y(o)[q * 90 + k] with the physical envelope
// env = (mu / 4 pi h) |o|^-(2 + order(k) + order(q)).
// Matrix products run as BF16 MFMA (v_mfma_f32_32x32x16_bf16 in the tiled
// kernels, v_mfma_f32_32x32x8bf16_1k in the generic one) with FP32
// accumulation; master weights and Adam moments are FP32.
//
// Training data are exact local expansions: a sample is (offset o, source
// cell s); its target is the L2 projection onto the 10-term quadratic basis
// of the exact non-singular stress of the cell's segments at the 27 tensor
// Gauss points of the cell at s - o, L_exact (6 x 10).  The loss is
// sum_s |W(o_s) M_s - L_exact_s|^2 / (den_s ns), with the envelope inside W.
#include "mlp.hpp"

#include <algorithm>
#include <type_traits>

namespace ddd {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);  // inf/nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

__device__ __forceinline__ float silu(float z) { return z / (1.f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = 1.f / (1.f + __expf(-z));
  return s * (1.f + z * (1.f - s));
}

// The layer epilogue on element (i, j) of C (row stride ldc): c = C value
// after the product, returned as stored; H written if given.
__device__ __forceinline__ float epilogue(float c, const float* __restrict__ bias, int act,
                                          unsigned short* __restrict__ H, int64_t e, int j) {
  if (!bias) return c;
  const float z = c + bias[j];
  if (H) H[e] = f2bf(act ? silu(z) : z);
  return z;
}

// C[M x N] (fp32, row-major, ldc) = sum_k A(i, k) B(k, j), A and B bf16 with
// element strides (a_i, a_k) and (b_k, b_j).  Workgroup: 4 wavefronts, a
// 64 x 64 tile of C, 32 x 32 per wavefront; K staged through LDS 32 at a time.
constexpr int GT = 64, GK = 32;

__global__ __launch_bounds__(256) void k_gemm_bf16(const unsigned short* __restrict__ A, int64_t a_i,
                                                   int64_t a_k, const unsigned short* __restrict__ B,
                                                   int64_t b_k, int64_t b_j, float* __restrict__ C,
                                                   int64_t ldc, int M, int N, int K, int accumulate,
                                                   const float* __restrict__ bias, int act,
                                                   unsigned short* __restrict__ H) {
  __shared__ unsigned short sA[GT][GK + 4];
  __shared__ unsigned short sB[GK][GT + 4];
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  const int i0 = blockIdx.y * GT, j0 = blockIdx.x * GT;
  const int wi = (wave / 2) * 32, wj = (wave % 2) * 32;
  f32x16 acc;
#pragma unroll
  for (int v = 0; v < 16; ++v) acc[v] = 0.f;
  for (int k0 = 0; k0 < K; k0 += GK) {
    for (int e = tid; e < GT * GK; e += 256) {
      const int r = e / GK, kk = e % GK;
      const int gi = i0 + r, gk = k0 + kk;
      sA[r][kk] = (gi < M && gk < K) ? A[gi * a_i + gk * a_k] : (unsigned short)0;
    }
    for (int e = tid; e < GK * GT; e += 256) {
      const int kk = e / GT, c = e % GT;
      const int gk = k0 + kk, gj = j0 + c;
      sB[kk][c] = (gk < K && gj < N) ? B[gk * b_k + gj * b_j] : (unsigned short)0;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < GK; ks += 8) {
      // A operand: row (lane % 32), k = ks + 4 (lane / 32) + 0..3
      // B operand: column (lane % 32), k = ks + 4 (lane / 32) + 0..3
      const int kb = ks + 4 * (lane / 32);
      bf16x4 a, b;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        a[q] = (short)sA[wi + lane % 32][kb + q];
        b[q] = (short)sB[kb + q][wj + lane % 32];
      }
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  // D layout: lane holds column (lane % 32), rows 8 (v / 4) + 4 (lane / 32) + v % 4.
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int gi = i0 + wi + 8 * (v / 4) + 4 * (lane / 32) + (v % 4);
    const int gj = j0 + wj + lane % 32;
    if (gi < M && gj < N) {
      const int64_t e = (int64_t)gi * ldc + gj;
      C[e] = epilogue(accumulate ? C[e] + acc[v] : acc[v], bias, act, H, e, gj);
    }
  }
}

// Faster product for aligned operands: 128 x 128 tile per workgroup of 4
// wavefronts (64 x 64 each, 2 x 2 blocks of v_mfma_f32_32x32x16_bf16), K
// staged 64 at a time.  Both operands sit in LDS with k contiguous (A as
// [row][k], B as [col][k]), so each lane fetches its 8 k-values with one
// 16-byte read; global loads are 8-element vectors along whichever index is
// contiguous, and the next two K tiles are loaded into registers while the
// current one is multiplied.  Requires 16-byte aligned bases, the non-unit
// stride of the contiguous operand a multiple of 8, and the contiguous
// extent of each operand a multiple of 8 (k, or M / N).
//
// Split K: workgroup z of gridDim.z takes k in [z kChunk, (z + 1) kChunk) and
// writes its partial product to C + z cStride (accumulate must be 0); a
// second kernel adds the partials in z order.  AK / BKC (which operands are
// contiguous along k) and EPI (the epilogue) are as for k_gemm_bf16_t256
// below; two wavefronts a SIMD.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
constexpr int BT = 128, BK = 64, LDK = BK + 8;

// LDS column of k in row `row` of a staged tile: the eight 16-byte k blocks
// of a row are permuted by an XOR with (row / 8) mod 8.  An 8-element block
// stays contiguous and aligned, so every MFMA operand read is still one
// 16-byte read; the transposing stash, whose 16 lanes sharing a k write rows
// 8 apart (one LDS bank without the permutation), then spreads over eight.
__device__ __forceinline__ int swz(int row, int k) { return k ^ (((row >> 3) & 7) << 3); }

struct TileRegs {
  u16x8 a[4], b[4];
  unsigned okA, okB;  // bit h: vector h lies inside the operand
};

template <bool AK, bool BKC, int EPI>
__global__ __launch_bounds__(256, 2) void k_gemm_bf16_t128(const unsigned short* __restrict__ A,
                                                        int64_t a_i, int64_t a_k,
                                                        const unsigned short* __restrict__ B,
                                                        int64_t b_k, int64_t b_j,
                                                        float* __restrict__ C, int64_t ldc, int M,
                                                        int N, int K, int accumulate, int kChunk,
                                                        int64_t cStride,
                                                        const float* __restrict__ bias, int act,
                                                        unsigned short* __restrict__ H) {
  __shared__ __attribute__((aligned(16))) unsigned short sA[BT][LDK];
  __shared__ __attribute__((aligned(16))) unsigned short sB[BT][LDK];
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  const int i0 = blockIdx.y * BT, j0 = blockIdx.x * BT;
  const int wi = (wave / 2) * 64, wj = (wave % 2) * 64;
  const int kBeg = blockIdx.z * kChunk;
  const int kEnd = min(K, kBeg + kChunk);
  C += blockIdx.z * cStride;
  // the bias of the lane's two columns, loaded now so that the epilogue
  // does not wait for it
  float bj[2] = {0.f, 0.f};
  if constexpr (EPI >= 2) {
#pragma unroll
    for (int n = 0; n < 2; ++n) bj[n] = bias[min(j0 + wj + 32 * n + lane % 32, N - 1)];
  }
  f32x16 acc[2][2];
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int n = 0; n < 2; ++n)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[m][n][v] = 0.f;

  // Vector v (0..1023) of the 128 x 64 A tile and of the 64 x 128 B tile.
  // The launcher guarantees that the contiguous extent of each operand is a
  // multiple of 8, so a vector lies wholly inside or wholly outside it:
  // outside ones read a clamped in-range address and are replaced by zeros
  // when written to LDS, so that nothing waits for the loads before then.
  auto fetch = [&](int k0, TileRegs& R) {
    R.okA = R.okB = 0;
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const int v = tid + 256 * h;
      if constexpr (AK) {  // 8 consecutive k of one row
        const int gi = i0 + v / 8, gk = k0 + (v % 8) * 8;
        R.okA |= (unsigned)(gi < M && gk < kEnd) << h;
        R.a[h] = *(const u16x8*)(A + (int64_t)min(gi, M - 1) * a_i + min(gk, K - 8));
      } else {  // 8 consecutive rows at one k
        const int kk = v / 16, rq = (v % 16) * 8;
        const int gi = i0 + rq, gk = k0 + kk;
        R.okA |= (unsigned)(gk < kEnd && gi < M) << h;
        R.a[h] = *(const u16x8*)(A + min(gi, M - 8) + (int64_t)min(gk, K - 1) * a_k);
      }
      if constexpr (BKC) {  // 8 consecutive k of one column
        const int gj = j0 + v / 8, gk = k0 + (v % 8) * 8;
        R.okB |= (unsigned)(gj < N && gk < kEnd) << h;
        R.b[h] = *(const u16x8*)(B + min(gk, K - 8) + (int64_t)min(gj, N - 1) * b_j);
      } else {  // 8 consecutive columns at one k
        const int kk = v / 16, cq = (v % 16) * 8;
        const int gj = j0 + cq, gk = k0 + kk;
        R.okB |= (unsigned)(gk < kEnd && gj < N) << h;
        R.b[h] = *(const u16x8*)(B + (int64_t)min(gk, K - 1) * b_k + min(gj, N - 8));
      }
    }
  };
  auto stash = [&](const TileRegs& R) {
    const u16x8 zero8 = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const int v = tid + 256 * h;
      const u16x8 ta = (R.okA >> h) & 1 ? R.a[h] : zero8;
      const u16x8 tb = (R.okB >> h) & 1 ? R.b[h] : zero8;
      if constexpr (AK) {
        const int r = v / 8;
        *(u16x8*)&sA[r][swz(r, (v % 8) * 8)] = ta;
      } else {
        const int kk = v / 16, rq = (v % 16) * 8;
#pragma unroll
        for (int q = 0; q < 8; ++q) sA[rq + q][swz(rq + q, kk)] = ta[q];
      }
      if constexpr (BKC) {
        const int c = v / 8;
        *(u16x8*)&sB[c][swz(c, (v % 8) * 8)] = tb;
      } else {
        const int kk = v / 16, cq = (v % 16) * 8;
#pragma unroll
        for (int q = 0; q < 8; ++q) sB[cq + q][swz(cq + q, kk)] = tb[q];
      }
    }
  };

  auto products = [&] {
#pragma unroll
    for (int ks = 0; ks < BK; ks += 16) {
      const int kb = ks + 8 * (lane / 32);
      bf16x8 a[2], b[2];
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        const int r = wi + 32 * m + lane % 32;
        a[m] = *(const bf16x8*)&sA[r][swz(r, kb)];
      }
#pragma unroll
      for (int n = 0; n < 2; ++n) {
        const int c = wj + 32 * n + lane % 32;
        b[n] = *(const bf16x8*)&sB[c][swz(c, kb)];
      }
#pragma unroll
      for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int n = 0; n < 2; ++n)
          acc[m][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[m], b[n], acc[m][n], 0, 0, 0);
    }
  };
  // Two K tiles in flight: while tile t is multiplied from LDS, tile t + 1
  // waits in one register set and tile t + 2 is loaded into the other (one
  // step of lookahead leaves these small products waiting on memory
  // latency).
  TileRegs R0, R1;
  if (kBeg < kEnd) {
    fetch(kBeg, R0);
    stash(R0);
  }
  if (kBeg + BK < kEnd) fetch(kBeg + BK, R1);
  __syncthreads();
  for (int k0 = kBeg; k0 < kEnd; k0 += 2 * BK) {
    if (k0 + 2 * BK < kEnd) fetch(k0 + 2 * BK, R0);
    products();
    __syncthreads();
    if (k0 + BK >= kEnd) break;
    stash(R1);
    __syncthreads();
    if (k0 + 3 * BK < kEnd) fetch(k0 + 3 * BK, R1);
    products();
    __syncthreads();
    if (k0 + 2 * BK < kEnd) stash(R0);
    __syncthreads();
  }
  // Epilogue (as epilogue() above, with the bias preloaded): the lane holds
  // column wj + 32 n + lane % 32 of rows wi + 32 m + 4 (lane / 32) +
  // 8 (v / 4) + v % 4.
  const int ld32 = (int)ldc;
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    const int gi0 = i0 + wi + 32 * m + 4 * (lane / 32);
#pragma unroll
    for (int n = 0; n < 2; ++n) {
      const int gj = j0 + wj + 32 * n + lane % 32;
      if (gj >= N) continue;
      float* cb = C + (int64_t)gi0 * ldc + gj;
      unsigned short* hb = EPI == 3 ? H + (int64_t)gi0 * ldc + gj : nullptr;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int dr = 8 * (v / 4) + (v % 4);
        if (gi0 + dr >= M) continue;
        const int off = dr * ld32;
        float x = EPI == 1 ? cb[off] + acc[m][n][v] : acc[m][n][v];
        if constexpr (EPI >= 2) x = x + bj[n];
        if constexpr (EPI == 3) hb[off] = f2bf(act ? silu(x) : x);
        cb[off] = x;
      }
    }
  }
}

// The large products: a 256 x 256 tile per workgroup of 4 wavefronts, each
// wavefront 128 x 128 (4 x 4 blocks of v_mfma_f32_32x32x16_bf16; its 256
// accumulator registers fill the AGPR half of the register file, one
// wavefront per SIMD).  A 128 x 128 wavefront tile reads 8 KB of operands
// from LDS per 16 MFMAs, half the LDS traffic per flop of a 64 x 64 one,
// whose operand reads alone would match the LDS bandwidth at the MFMA rate.
// K is staged 64 at a time through two LDS buffers: the next K tile is
// loaded into registers while the current one is multiplied, then written
// to the other buffer, with one barrier per K step.  AK / BKC: the operand is
// contiguous along k (else along i / j, transposed on the way into LDS).
// Same layout, alignment and split-K conventions as k_gemm_bf16_t128.
// EPI: 0 stores the product, 1 adds it to C, 2 adds the bias, 3 adds the
// bias and writes the activation to H (the unrolled activation code of 256
// values is kept out of the other variants: it would not fit the
// instruction cache next to the main loop).
constexpr int HT = 256;

template <bool AK, bool BKC, int EPI>
__global__ __launch_bounds__(256, 1) void k_gemm_bf16_t256(const unsigned short* __restrict__ A,
                                                           int64_t a_i, int64_t a_k,
                                                           const unsigned short* __restrict__ B,
                                                           int64_t b_k, int64_t b_j,
                                                           float* __restrict__ C, int64_t ldc,
                                                           int M, int N, int K, int accumulate,
                                                           int kChunk, int64_t cStride,
                                                           const float* __restrict__ bias, int act,
                                                           unsigned short* __restrict__ H) {
  __shared__ __attribute__((aligned(16))) unsigned short sA[2][HT][LDK];
  __shared__ __attribute__((aligned(16))) unsigned short sB[2][HT][LDK];
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  const int i0 = blockIdx.y * HT, j0 = blockIdx.x * HT;
  const int wi = (wave / 2) * 128, wj = (wave % 2) * 128;
  const int kBeg = blockIdx.z * kChunk;
  const int kEnd = min(K, kBeg + kChunk);
  C += blockIdx.z * cStride;
  // the bias of the lane's four columns, loaded now so that the epilogue
  // does not wait for it
  float bj[4] = {0.f, 0.f, 0.f, 0.f};
  if constexpr (EPI >= 2) {
#pragma unroll
    for (int n = 0; n < 4; ++n) bj[n] = bias[min(j0 + wj + 32 * n + lane % 32, N - 1)];
  }
  f32x16 acc[4][4];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int n = 0; n < 4; ++n)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[m][n][v] = 0.f;

  // Vector v = tid + 256 h (h < 8) of the 256 x 64 operand tiles: k
  // contiguous, row v / 8 and k (v % 8) * 8; else k v / 32 and rows
  // (v % 32) * 8 .. + 7.  The launcher guarantees that the contiguous
  // extent of each operand is a multiple of 8, so a vector lies wholly
  // inside or wholly outside it: outside ones read a clamped in-range
  // address, and are replaced by zeros when written to LDS (bit h of okA /
  // okB), so that nothing waits for the loads before the products.
  u16x8 ra[8], rb[8];
  unsigned okA = 0, okB = 0;
  auto fetch = [&](int k0) {
    okA = okB = 0;
#pragma unroll
    for (int h = 0; h < 8; ++h) {
      const int v = tid + 256 * h;
      if constexpr (AK) {
        const int gi = i0 + v / 8, gk = k0 + (v % 8) * 8;
        okA |= (unsigned)(gi < M && gk < kEnd) << h;
        ra[h] = *(const u16x8*)(A + (int64_t)min(gi, M - 1) * a_i + min(gk, K - 8));
      } else {
        const int gk = k0 + v / 32, gi = i0 + (v % 32) * 8;
        okA |= (unsigned)(gk < kEnd && gi < M) << h;
        ra[h] = *(const u16x8*)(A + min(gi, M - 8) + (int64_t)min(gk, K - 1) * a_k);
      }
      if constexpr (BKC) {
        const int gj = j0 + v / 8, gk = k0 + (v % 8) * 8;
        okB |= (unsigned)(gj < N && gk < kEnd) << h;
        rb[h] = *(const u16x8*)(B + min(gk, K - 8) + (int64_t)min(gj, N - 1) * b_j);
      } else {
        const int gk = k0 + v / 32, gj = j0 + (v % 32) * 8;
        okB |= (unsigned)(gk < kEnd && gj < N) << h;
        rb[h] = *(const u16x8*)(B + (int64_t)min(gk, K - 1) * b_k + min(gj, N - 8));
      }
    }
  };
  auto stash = [&](int buf) {
    const u16x8 zero8 = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int h = 0; h < 8; ++h) {
      const int v = tid + 256 * h;
      const u16x8 ta = (okA >> h) & 1 ? ra[h] : zero8;
      const u16x8 tb = (okB >> h) & 1 ? rb[h] : zero8;
      if constexpr (AK) {
        const int r = v / 8;
        *(u16x8*)&sA[buf][r][swz(r, (v % 8) * 8)] = ta;
      } else {
        const int kk = v / 32, rq = (v % 32) * 8;
#pragma unroll
        for (int q = 0; q < 8; ++q) sA[buf][rq + q][swz(rq + q, kk)] = ta[q];
      }
      if constexpr (BKC) {
        const int c = v / 8;
        *(u16x8*)&sB[buf][c][swz(c, (v % 8) * 8)] = tb;
      } else {
        const int kk = v / 32, cq = (v % 32) * 8;
#pragma unroll
        for (int q = 0; q < 8; ++q) sB[buf][cq + q][swz(cq + q, kk)] = tb[q];
      }
    }
  };

  int cur = 0;
  if (kBeg < kEnd) {
    fetch(kBeg);
    stash(0);
  }
  __syncthreads();
  for (int k0 = kBeg; k0 < kEnd; k0 += BK) {
    const bool more = k0 + BK < kEnd;
    if (more) fetch(k0 + BK);  // in flight during the products below
    // Operands of k sub-step ks + 16 are read from LDS while the 16
    // products of sub-step ks run (two register sets).
    bf16x8 a[2][4], b[2][4];
    auto operands = [&](int ks, int set) {
      const int kb = ks + 8 * (lane / 32);
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int r = wi + 32 * m + lane % 32;
        a[set][m] = *(const bf16x8*)&sA[cur][r][swz(r, kb)];
      }
#pragma unroll
      for (int n = 0; n < 4; ++n) {
        const int c = wj + 32 * n + lane % 32;
        b[set][n] = *(const bf16x8*)&sB[cur][c][swz(c, kb)];
      }
    };
    operands(0, 0);
#pragma unroll
    for (int q = 0; q < BK / 16; ++q) {
      if (q + 1 < BK / 16) operands(16 * (q + 1), (q + 1) & 1);
      // keeps the reads ahead of the products (the scheduler would
      // otherwise reuse one register set and wait on every read)
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int n = 0; n < 4; ++n)
          acc[m][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[q & 1][m], b[q & 1][n], acc[m][n],
                                                              0, 0, 0);
    }
    // The other buffer was last read before the previous barrier.
    if (more) stash(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }
  // Epilogue (as epilogue() above): the lane holds column
  // wj + 32 n + lane % 32 of rows wi + 32 m + 4 (lane / 32) + 8 (v / 4) + v % 4;
  // per block one base address, per v a row offset common to all blocks.
  const int ld32 = (int)ldc;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int gi0 = i0 + wi + 32 * m + 4 * (lane / 32);
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const int gj = j0 + wj + 32 * n + lane % 32;
      if (gj >= N) continue;
      float* cb = C + (int64_t)gi0 * ldc + gj;
      unsigned short* hb = EPI == 3 ? H + (int64_t)gi0 * ldc + gj : nullptr;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int dr = 8 * (v / 4) + (v % 4);
        if (gi0 + dr >= M) continue;
        const int off = dr * ld32;
        float x = EPI == 1 ? cb[off] + acc[m][n][v] : acc[m][n][v];
        if constexpr (EPI >= 2) x = x + bj[n];
        if constexpr (EPI == 3) hb[off] = f2bf(act ? silu(x) : x);
        cb[off] = x;
      }
    }
  }
}

// C (+)= sum over z < nSplit of the partials P + z stride (M x N, ldc N),
// then the epilogue.
__global__ void k_splitk_reduce(const float* __restrict__ P, int64_t stride, int nSplit, int64_t n,
                                int N, float* __restrict__ C, int accumulate,
                                const float* __restrict__ bias, int act,
                                unsigned short* __restrict__ H) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n) return;
  float a = 0.f;
  for (int z = 0; z < nSplit; ++z) a += P[z * stride + e];
  C[e] = epilogue(accumulate ? C[e] + a : a, bias, act, H, e, (int)(e % N));
}

// Column sums of bf16 matrices in a fixed order.  A workgroup of 4
// wavefronts takes 64 columns (one a lane) of a chunk of kColRows rows;
// wavefront w sums rows w kColRows / 4 .. (w + 1) kColRows / 4 - 1 of the
// chunk in order, and the four sums are added in wavefront order.  ACT: the
// matrix is formed on the way, dZ = bf16(dH * SiLU'(Z)), and stored.
constexpr int kColWaveRows = kColRows / 4;

template <bool ACT>
__global__ __launch_bounds__(256) void k_colsum_part(const float* __restrict__ dH,
                                                     const float* __restrict__ Z,
                                                     unsigned short* __restrict__ dZ, int rows,
                                                     int cols, float* __restrict__ part) {
  __shared__ float sw[4][64];
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  const int c = blockIdx.x * 64 + lane;
  const int r0 = blockIdx.y * kColRows + w * kColWaveRows;
  const int r1 = min(rows, r0 + kColWaveRows);
  float a = 0.f;
  if (c < cols) {
#pragma unroll 8
    for (int r = r0; r < r1; ++r) {
      const int64_t e = (int64_t)r * cols + c;
      if constexpr (ACT) {
        const unsigned short d = f2bf(dH[e] * silu_grad(Z[e]));
        dZ[e] = d;
        a += bf2f(d);
      } else {
        a += bf2f(dZ[e]);
      }
    }
  }
  sw[w][lane] = a;
  __syncthreads();
  if (w == 0 && c < cols)
    part[(int64_t)blockIdx.y * cols + c] = ((sw[0][lane] + sw[1][lane]) + sw[2][lane]) + sw[3][lane];
}

// db[c] = sum of the chunk partials: wavefront w adds chunks
// w q .. (w + 1) q - 1 (q = ceil(nChunks / 4)) in order, then the four sums
// in wavefront order.
__global__ __launch_bounds__(256) void k_colsum_final(const float* __restrict__ part, int cols,
                                                      int nChunks, float* __restrict__ db) {
  __shared__ float sw[4][64];
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  const int c = blockIdx.x * 64 + lane;
  const int q = (nChunks + 3) / 4;
  const int h0 = w * q, h1 = min(nChunks, h0 + q);
  float a = 0.f;
  if (c < cols) {
#pragma unroll 8
    for (int h = h0; h < h1; ++h) a += part[(int64_t)h * cols + c];
  }
  sw[w][lane] = a;
  __syncthreads();
  if (w == 0 && c < cols) db[c] = ((sw[0][lane] + sw[1][lane]) + sw[2][lane]) + sw[3][lane];
}

__global__ void k_f32_to_bf16(const float* __restrict__ x, int64_t n, unsigned short* __restrict__ y) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e < n) y[e] = f2bf(x[e]);
}

// Offset features: 1/|o|, 1/|o|^2, 1/|o|^3, then the monomials
// ux^a uy^b uz^c of the direction u = o/|o| for degrees d = 1..7, in the
// order a = d..0, b = d-a..0, c = d-a-b, then 1/R_a, 1/R_a^2, 1/R_a^3 with
// R_a^2 = |o|^2 + aH^2 (the core radius in cell units); zero padded.
__device__ void offset_features(int ox, int oy, int oz, float aH, float* f) {
  for (int q = 0; q < kMlpIn; ++q) f[q] = 0.f;
  if (ox == 0 && oy == 0 && oz == 0) return;  // the zero offset is never used
  const float R2 = (float)(ox * ox + oy * oy + oz * oz);
  const float R = sqrtf(R2);
  const float ir = 1.f / R;
  const float u[3] = {ox * ir, oy * ir, oz * ir};
  f[0] = ir;
  f[1] = ir * ir;
  f[2] = ir * ir * ir;
  float px[kMonoDeg + 1], py[kMonoDeg + 1], pz[kMonoDeg + 1];
  px[0] = py[0] = pz[0] = 1.f;
  for (int e = 1; e <= kMonoDeg; ++e) {
    px[e] = px[e - 1] * u[0];
    py[e] = py[e - 1] * u[1];
    pz[e] = pz[e - 1] * u[2];
  }
  int q = 3;
  for (int d = 1; d <= kMonoDeg; ++d)
    for (int a = d; a >= 0; --a)
      for (int b = d - a; b >= 0; --b) f[q++] = px[a] * py[b] * pz[d - a - b];
  const float ira = 1.f / sqrtf(R2 + aH * aH);
  f[q] = ira;
  f[q + 1] = ira * ira;
  f[q + 2] = ira * ira * ira;
}

__global__ __launch_bounds__(256) void k_features(const int* __restrict__ off3, int ns, float aH,
                                                  unsigned short* __restrict__ X) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  float f[kMlpIn];
  offset_features(off3[3 * s], off3[3 * s + 1], off3[3 * s + 2], aH, f);
  for (int q = 0; q < kMlpIn; ++q) X[(int64_t)s * kMlpIn + q] = f2bf(f[q]);
}

__device__ __forceinline__ int moment_order(int k) { return k < 9 ? 0 : (k < 36 ? 1 : 2); }
__device__ __forceinline__ int poly_order(int l) { return l == 0 ? 0 : (l < 4 ? 1 : 2); }

// Per sample: predicted local expansion L_hat[q] = sum_k W(o)[q][k] M[k]
// (q = 10 c + l), loss term |L_hat - L|^2 / (den ns) and dY (bf16, the
// backward products' operand).  One workgroup of 256 per sample.
__global__ __launch_bounds__(256) void k_contract(const float* __restrict__ Y,
                                                  const int* __restrict__ off3,
                                                  const float* __restrict__ Ms,
                                                  const float* __restrict__ Lt,
                                                  const float* __restrict__ den, int ns,
                                                  float envScale,
                                                  unsigned short* __restrict__ dY,
                                                  float* __restrict__ lossPer) {
  const int s = blockIdx.x;
  const int tid = threadIdx.x;
  __shared__ float sM[kMoments];
  __shared__ float envK[kMoments];
  __shared__ float envL[kPoly];
  __shared__ float r[kLocal];
  const int ox = off3[3 * s], oy = off3[3 * s + 1], oz = off3[3 * s + 2];
  const float iR = 1.f / sqrtf((float)(ox * ox + oy * oy + oz * oz));
  const float* y = Y + (int64_t)s * kMlpOutPad;
  if (tid < kMoments) {
    sM[tid] = Ms[(int64_t)s * kMoments + tid];
    envK[tid] = powf(iR, (float)moment_order(tid));
  }
  if (tid < kPoly) envL[tid] = envScale * powf(iR, (float)(2 + poly_order(tid)));
  __syncthreads();
  if (tid < kLocal) {
    const int l = tid % kPoly;
    float a = 0.f;
    for (int k = 0; k < kMoments; ++k) a += envK[k] * y[tid * kMoments + k] * sM[k];
    r[tid] = envL[l] * a - Lt[(int64_t)s * kLocal + tid];
  }
  __syncthreads();
  // dY[q][k] = A[q] B[k], A = w r envL, B = envK M (the padding zero)
  const float w = 2.f / (den[s] * ns);
  __shared__ float A[kLocal], B[kMoments];
  if (tid < kLocal) A[tid] = w * r[tid] * envL[tid % kPoly];
  if (tid < kMoments) B[tid] = envK[tid] * sM[tid];
  __syncthreads();
  // two k a thread (one 32-bit store), 45 pairs x 5 rows a step
  constexpr int kPairs = kMoments / 2, kRows = 256 / kPairs;
  static_assert(kMoments % 2 == 0 && (kMlpOutPad - kChan) % 2 == 0 &&
                    kPairs * kRows + (kMlpOutPad - kChan) / 2 <= 256,
                "one workgroup of 256 writes a sample's dY in 32-bit words");
  unsigned* d32 = reinterpret_cast<unsigned*>(dY + (int64_t)s * kMlpOutPad);
  if (tid < kPairs * kRows) {
    const int kp = tid % kPairs;
    const float b0 = B[2 * kp], b1 = B[2 * kp + 1];
    for (int q = tid / kPairs; q < kLocal; q += kRows) {
      const float a = A[q];
      d32[q * kPairs + kp] = (unsigned)f2bf(a * b0) | ((unsigned)f2bf(a * b1) << 16);
    }
  } else if (tid < kPairs * kRows + (kMlpOutPad - kChan) / 2) {
    d32[kChan / 2 + (tid - kPairs * kRows)] = 0u;
  }
  if (tid == 0) {
    float a = 0.f;
    for (int q = 0; q < kLocal; ++q) a += r[q] * r[q];
    lossPer[s] = a / (den[s] * ns);
  }
}

// W(o) = env * y for the offsets of the grid, [o][60][90] fp32.
__global__ void k_operator_from_y(const float* __restrict__ Y, const int* __restrict__ off3, int no,
                                  float envScale, float* __restrict__ W) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= (int64_t)no * kChan) return;
  const int o = (int)(e / kChan);
  const int ch = (int)(e % kChan);
  const int ox = off3[3 * o], oy = off3[3 * o + 1], oz = off3[3 * o + 2];
  const int q = ch / kMoments, k = ch % kMoments;
  const int l = q % kPoly;
  float v = 0.f;
  if (ox != 0 || oy != 0 || oz != 0) {
    const float iR = 1.f / sqrtf((float)(ox * ox + oy * oy + oz * oz));
    v = envScale * powf(iR, (float)(2 + moment_order(k) + poly_order(l))) * Y[(int64_t)o * kMlpOutPad + ch];
  }
  W[e] = v;
}

// One Adam step; pb receives the bf16 copy of the updated parameters.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float c1,
                       float c2, float eps, unsigned short* __restrict__ pb) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gg = g[e];
  const float mm = b1 * m[e] + (1.f - b1) * gg;
  const float vv = b2 * v[e] + (1.f - b2) * gg * gg;
  m[e] = mm;
  v[e] = vv;
  const float q = p[e] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
  p[e] = q;
  pb[e] = f2bf(q);
}

// Batch b of the training set: sample idx[b]; moments from the cell array
// M (feature-major, nc cells).  cubic: the sample is carried into the frame
// of its canonical offset o' (o = R o'): offset o', moments T_M(R)^T M and
// target det(R) T_L(R)^T L, so that the network learns W at canonical
// offsets only.
__global__ void k_gather_batch(const int* __restrict__ idx, int nb, const int* __restrict__ off3,
                               const int* __restrict__ src, const float* __restrict__ M, int nc,
                               const float* __restrict__ Lt, const float* __restrict__ den,
                               int cubic, int* __restrict__ bOff, float* __restrict__ bM,
                               float* __restrict__ bL, float* __restrict__ bDen) {
  const int b = blockIdx.x;
  const int t = threadIdx.x;
  if (b >= nb) return;
  const int64_t s = idx[b];
  const int c = src[s];
  AxisMap R{{0, 1, 2}, {1, 1, 1}};
  int oc[3] = {off3[s * 3], off3[s * 3 + 1], off3[s * 3 + 2]};
  if (cubic) canonical_offset(oc[0], oc[1], oc[2], oc, R);
  if (t < kMoments) {
    int k2 = t, sg = 1;
    if (cubic) moment_map(R, t, k2, sg);
    bM[(int64_t)b * kMoments + t] = (float)sg * M[(int64_t)k2 * nc + c];
  }
  if (t < kLocal) {
    int q2 = t, sg = 1;
    if (cubic) {
      local_map(R, t, q2, sg);
      sg *= axis_det(R);
    }
    bL[(int64_t)b * kLocal + t] = (float)sg * Lt[s * kLocal + q2];
  }
  if (t < 3) bOff[(int64_t)b * 3 + t] = oc[t];
  if (t == 0) bDen[b] = den[s];
}

// W(o) for every offset o of the (2G - 1)^3 grid (x fastest) from Wc, the
// operator at the canonical offsets (canonical_index order):
// W(o) = det(R) T_L(R) W(o') T_M(R)^T for o = R o', averaged over the maps that
// fix o' (offsets with equal or zero components), so that the installed
// operator is exactly equivariant.  One workgroup per offset.
__global__ __launch_bounds__(256) void k_operator_expand(const float* __restrict__ Wc, int G,
                                                         float* __restrict__ W) {
  const int S = 2 * G - 1;
  const int64_t o = blockIdx.x;
  __shared__ AxisMap maps[48];
  __shared__ int nmap, ci;
  if (threadIdx.x == 0) {
    const int ox = (int)(o % S) - (G - 1), oy = (int)((o / S) % S) - (G - 1),
              oz = (int)(o / ((int64_t)S * S)) - (G - 1);
    int oc[3];
    AxisMap R;
    canonical_offset(ox, oy, oz, oc, R);
    int n = 0;
    for (int k = 0; k < 48; ++k) {
      const AxisMap Sk = axis_map(k);
      int v[3];
      for (int e = 0; e < 3; ++e) v[Sk.perm[e]] = Sk.sgn[e] * oc[e];
      if (v[0] == oc[0] && v[1] == oc[1] && v[2] == oc[2]) maps[n++] = inverse(compose(R, Sk));
    }
    nmap = n;
    ci = canonical_index(oc[0], oc[1], oc[2]);
  }
  __syncthreads();
  const float* wc = Wc + (int64_t)ci * kChan;
  const float inv = 1.f / (float)nmap;
  for (int e = threadIdx.x; e < kChan; e += blockDim.x) {
    const int Q = e / kMoments, K = e % kMoments;
    float a = 0.f;
    for (int m = 0; m < nmap; ++m) {
      int q, sq, k, sk;
      local_map(maps[m], Q, q, sq);
      moment_map(maps[m], K, k, sk);
      a += (float)(axis_det(maps[m]) * sq * sk) * wc[q * kMoments + k];
    }
    W[o * kChan + e] = a * inv;
  }
}

// Exact local expansion of source cell src[s] over the cell at offset
// -off3[s]: FP64 stress (the non-singular field of DislocationFEM.jl:142-299)
// at the 27 Gauss points, projected with Pj (10 x 27).  Thread per sample.
__global__ __launch_bounds__(64) void k_local_targets(const double* __restrict__ seg, int64_t ld, GridDev g,
                                const int* __restrict__ start, const int* __restrict__ order,
                                const int* __restrict__ src, const int* __restrict__ off3, int ns,
                                SigK<double> K, Projection Pj, float* __restrict__ Lt) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  const int c = src[s];
  const int cx = c % g.G, cy = (c / g.G) % g.G, cz = c / (g.G * g.G);
  const double tc[3] = {g.lo[0] + (cx - off3[3 * s] + 0.5) * g.h,
                        g.lo[1] + (cy - off3[3 * s + 1] + 0.5) * g.h,
                        g.lo[2] + (cz - off3[3 * s + 2] + 0.5) * g.h};
  double L[kLocal];
#pragma unroll
  for (int q = 0; q < kLocal; ++q) L[q] = 0.0;
  for (int p = 0; p < 27; ++p) {
    const v3<double> x = V(tc[0] + g.h * Pj.xi[p % 3], tc[1] + g.h * Pj.xi[(p / 3) % 3],
                           tc[2] + g.h * Pj.xi[p / 9]);
    double sg[6] = {0, 0, 0, 0, 0, 0};
    for (int e = start[c]; e < start[c + 1]; ++e) {
      const int i = order[e];
      const v3<double> n1 = V(seg[F_N1X * ld + i], seg[F_N1Y * ld + i], seg[F_N1Z * ld + i]);
      const v3<double> n2 = V(seg[F_N2X * ld + i], seg[F_N2Y * ld + i], seg[F_N2Z * ld + i]);
      const v3<double> b = V(seg[F_BX * ld + i], seg[F_BY * ld + i], seg[F_BZ * ld + i]);
      const v3<double> t = V(seg[F_TX * ld + i], seg[F_TY * ld + i], seg[F_TZ * ld + i]);
      seg_stress<double>(K, x - n1, n2 - n1, t, b, cross(t, b), sg);
    }
#pragma unroll
    for (int l = 0; l < kPoly; ++l) {
      const double w = Pj.P[l][p];
#pragma unroll
      for (int cc = 0; cc < 6; ++cc) L[cc * kPoly + l] += w * sg[cc];
    }
  }
#pragma unroll
  for (int q = 0; q < kLocal; ++q) Lt[(int64_t)s * kLocal + q] = (float)L[q];
}

}  // namespace

static unsigned blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

void mlp_gemm(const unsigned short* A, int64_t a_i, int64_t a_k, const unsigned short* B,
              int64_t b_k, int64_t b_j, float* C, int64_t ldc, int M, int N, int K, int accumulate,
              hipStream_t st, float* work, int64_t workFloats, MlpEpilogue ep) {
  const bool alignedA = ((uintptr_t)A % 16 == 0) &&
                        (a_k == 1 ? a_i % 8 == 0 : (a_i == 1 && a_k % 8 == 0));
  const bool alignedB = ((uintptr_t)B % 16 == 0) &&
                        (b_k == 1 ? b_j % 8 == 0 : (b_j == 1 && b_k % 8 == 0));
  // Each operand's contiguous extent a multiple of 8 (k for k-contiguous
  // operands; M or N otherwise).
  const bool whole8 = K % 8 == 0 && (a_k == 1 || M % 8 == 0) && (b_k == 1 || N % 8 == 0);
  if (alignedA && alignedB && whole8 && (int64_t)M * N * K >= (int64_t(1) << 33)) {
    // The large products on 256 x 256 tiles; fewer tiles than CUs: split K
    // into at most as many parts as fill the chip once.
    const int64_t tiles = (int64_t)blocks(N, HT) * blocks(M, HT);
    int split = 1;
    if (tiles < 256 && K >= 8 * BK && ldc == N) {
      split = (int)std::min<int64_t>(256 / tiles, K / (4 * BK));
      while (split > 1 && (int64_t)split * M * N > workFloats) --split;
    }
    int kChunk = K;
    if (split > 1) {
      kChunk = ((K + split - 1) / split + BK - 1) / BK * BK;
      split = (K + kChunk - 1) / kChunk;
    }
    const bool ak = (a_k == 1), bk = (b_k == 1);
    // epilogue variant: partials and plain products store, accumulate adds,
    // a bias applies the layer epilogue (bias with accumulate: the general
    // kernel below)
    const int epi = split > 1 ? 0 : (ep.bias ? (ep.H ? 3 : 2) : (accumulate ? 1 : 0));
    if (!(epi >= 2 && accumulate)) {
      auto pick = [&](auto e) {
        constexpr int E = decltype(e)::value;
        return ak ? (bk ? k_gemm_bf16_t256<true, true, E> : k_gemm_bf16_t256<true, false, E>)
                  : (bk ? k_gemm_bf16_t256<false, true, E> : k_gemm_bf16_t256<false, false, E>);
      };
      auto kern = epi == 0   ? pick(std::integral_constant<int, 0>{})
                  : epi == 1 ? pick(std::integral_constant<int, 1>{})
                  : epi == 2 ? pick(std::integral_constant<int, 2>{})
                             : pick(std::integral_constant<int, 3>{});
      dim3 grid(blocks(N, HT), blocks(M, HT), split);
      if (split > 1) {
        hipLaunchKernelGGL(kern, grid, dim3(256), 0, st, A, a_i, a_k, B, b_k, b_j, work, (int64_t)N,
                           M, N, K, 0, kChunk, (int64_t)M * N, nullptr, 0, nullptr);
        const int64_t n = (int64_t)M * N;
        hipLaunchKernelGGL(k_splitk_reduce, dim3(blocks(n, 256)), dim3(256), 0, st, work, n, split,
                           n, N, C, accumulate, ep.bias, ep.act, ep.H);
      } else {
        hipLaunchKernelGGL(kern, grid, dim3(256), 0, st, A, a_i, a_k, B, b_k, b_j, C, ldc, M, N, K,
                           accumulate, K, (int64_t)0, ep.bias, ep.act, ep.H);
      }
      return;
    }
  }
  if (alignedA && alignedB && whole8) {
    const bool ak = (a_k == 1), bk = (b_k == 1);
    auto pick128 = [&](auto e) {
      constexpr int E = decltype(e)::value;
      return ak ? (bk ? k_gemm_bf16_t128<true, true, E> : k_gemm_bf16_t128<true, false, E>)
                : (bk ? k_gemm_bf16_t128<false, true, E> : k_gemm_bf16_t128<false, false, E>);
    };
    const int64_t tiles = (int64_t)blocks(N, BT) * blocks(M, BT);
    // Few output tiles and a long K (the weight gradients): split K so that
    // the grid has about one workgroup a CU, if the workspace holds the
    // partials.  (Measured at 4,096 samples: 512 x 512 x 4,096 takes 113 us
    // unsplit, 23 us in 16 parts, 26 us in 32; the products with K = 512
    // are fastest unsplit.)
    int split = 1;
    if (tiles < 256 && K >= 2048 && ldc == N) {
      split = (int)std::min<int64_t>((256 + tiles - 1) / tiles, K / (4 * BK));
      while (split > 1 && (int64_t)split * M * N > workFloats) --split;
    }
    if (split > 1) {
      const int kChunk = ((K + split - 1) / split + BK - 1) / BK * BK;
      split = (K + kChunk - 1) / kChunk;
      dim3 grid(blocks(N, BT), blocks(M, BT), split);
      hipLaunchKernelGGL(pick128(std::integral_constant<int, 0>{}), grid, dim3(256), 0, st, A, a_i,
                         a_k, B, b_k, b_j, work, (int64_t)N, M, N, K, 0, kChunk, (int64_t)M * N,
                         nullptr, 0, nullptr);
      const int64_t n = (int64_t)M * N;
      hipLaunchKernelGGL(k_splitk_reduce, dim3(blocks(n, 256)), dim3(256), 0, st, work, n, split, n,
                         N, C, accumulate, ep.bias, ep.act, ep.H);
      return;
    }
    const int epi = ep.bias ? (ep.H ? 3 : 2) : (accumulate ? 1 : 0);
    if (!(epi >= 2 && accumulate)) {
      auto kern = epi == 0   ? pick128(std::integral_constant<int, 0>{})
                  : epi == 1 ? pick128(std::integral_constant<int, 1>{})
                  : epi == 2 ? pick128(std::integral_constant<int, 2>{})
                             : pick128(std::integral_constant<int, 3>{});
      dim3 grid(blocks(N, BT), blocks(M, BT), 1);
      hipLaunchKernelGGL(kern, grid, dim3(256), 0, st, A, a_i, a_k, B, b_k, b_j, C, ldc, M, N, K,
                         accumulate, K, (int64_t)0, ep.bias, ep.act, ep.H);
      return;
    }
  }
  dim3 grid(blocks(N, GT), blocks(M, GT));
  hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, st, A, a_i, a_k, B, b_k, b_j, C, ldc, M, N, K,
                     accumulate, ep.bias, ep.act, ep.H);
}

void mlp_colsum(const unsigned short* dZ, int rows, int cols, float* part, float* db,
                hipStream_t st) {
  const int nCh = (rows + kColRows - 1) / kColRows;
  hipLaunchKernelGGL(k_colsum_part<false>, dim3(blocks(cols, 64), nCh), dim3(256), 0, st, nullptr,
                     nullptr, const_cast<unsigned short*>(dZ), rows, cols, part);
  hipLaunchKernelGGL(k_colsum_final, dim3(blocks(cols, 64)), dim3(256), 0, st, part, cols, nCh, db);
}

void mlp_act_grad_colsum(const float* dH, const float* Z, int rows, int cols, unsigned short* dZ,
                         float* part, float* db, hipStream_t st) {
  const int nCh = (rows + kColRows - 1) / kColRows;
  hipLaunchKernelGGL(k_colsum_part<true>, dim3(blocks(cols, 64), nCh), dim3(256), 0, st, dH, Z, dZ,
                     rows, cols, part);
  hipLaunchKernelGGL(k_colsum_final, dim3(blocks(cols, 64)), dim3(256), 0, st, part, cols, nCh, db);
}

void mlp_to_bf16(const float* x, int64_t n, unsigned short* y, hipStream_t st) {
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(blocks(n, 256)), dim3(256), 0, st, x, n, y);
}

void mlp_features(const int* off3, int ns, float aH, unsigned short* X, hipStream_t st) {
  hipLaunchKernelGGL(k_features, dim3(blocks(ns, 256)), dim3(256), 0, st, off3, ns, aH, X);
}

void mlp_contract(const float* Y, const int* off3, const float* M, const float* Lt,
                  const float* den, int ns, float envScale, unsigned short* dY, float* lossPer,
                  hipStream_t st) {
  hipLaunchKernelGGL(k_contract, dim3(ns), dim3(256), 0, st, Y, off3, M, Lt, den, ns, envScale, dY,
                     lossPer);
}

void mlp_operator_from_y(const float* Y, const int* off3, int no, float envScale, float* W,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_operator_from_y, dim3(blocks((int64_t)no * kChan, 256)), dim3(256), 0, st, Y,
                     off3, no, envScale, W);
}

void mlp_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
              float c1, float c2, float eps, unsigned short* pb, hipStream_t st) {
  hipLaunchKernelGGL(k_adam, dim3(blocks(n, 256)), dim3(256), 0, st, p, g, m, v, n, lr, b1, b2, c1,
                     c2, eps, pb);
}

void gather_batch(const int* idx, int nb, const int* off3, const int* src, const float* M, int nc,
                  const float* Lt, const float* den, int cubic, int* bOff, float* bM, float* bL,
                  float* bDen, hipStream_t st) {
  hipLaunchKernelGGL(k_gather_batch, dim3(nb), dim3(128), 0, st, idx, nb, off3, src, M, nc, Lt, den,
                     cubic, bOff, bM, bL, bDen);
}

void operator_expand(const float* Wc, int G, float* W, hipStream_t st) {
  const int64_t S = 2 * G - 1;
  hipLaunchKernelGGL(k_operator_expand, dim3((unsigned)(S * S * S)), dim3(256), 0, st, Wc, G, W);
}

void local_targets(const double* seg, int64_t ld, const GridDev& g, const int* start,
                   const int* order, const int* src, const int* off3, int ns, const SigK<double>& K,
                   const Projection& Pj, float* Lt, hipStream_t st) {
  hipLaunchKernelGGL(k_local_targets, dim3(blocks(ns, 64)), dim3(64), 0, st, seg, ld, g, start, order,
                     src, off3, ns, K, Pj, Lt);
}

}  // namespace ddd
