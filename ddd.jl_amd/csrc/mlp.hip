// The far-field network (no reference counterpart): a 4-layer MLP mapping a
// cell offset o to the 60 x 90 operator W(o) of the hybrid far field,
//   phi(o) (32) -> 256 -> 256 -> 256 -> 5408 (5400 used), SiLU activations,
// W(o)[q][k] = env(o, q, k) * y(o)[q * 90 + k] with the physical envelope
// env = (mu / 4 pi h) |o|^-(2 + order(k) + order(q)).
// Matrix products run as BF16 MFMA (v_mfma_f32_32x32x8bf16_1k) with FP32
// accumulation; master weights and Adam moments are FP32.
//
// Training data are exact non-singular stresses: a sample is (offset, source
// cell moments M (90), target basis P(rho) (10), exact stress (6)); the loss
// is sum_s |sigma_hat_s - sigma_s|^2 / den_s with
// sigma_hat_s[c] = sum_{l,k} W(o_s)[c * 10 + l][k] M_s[k] P_s[l].
#include "mlp.hpp"

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

// C[M x N] (fp32, row-major, ldc) = sum_k A(i, k) B(k, j), A and B bf16 with
// element strides (a_i, a_k) and (b_k, b_j).  Workgroup: 4 wavefronts, a
// 64 x 64 tile of C, 32 x 32 per wavefront; K staged through LDS 32 at a time.
constexpr int GT = 64, GK = 32;

__global__ __launch_bounds__(256) void k_gemm_bf16(const unsigned short* __restrict__ A, int64_t a_i,
                                                   int64_t a_k, const unsigned short* __restrict__ B,
                                                   int64_t b_k, int64_t b_j, float* __restrict__ C,
                                                   int64_t ldc, int M, int N, int K, int accumulate) {
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
      float* p = C + (int64_t)gi * ldc + gj;
      *p = accumulate ? *p + acc[v] : acc[v];
    }
  }
}

__device__ __forceinline__ float silu(float z) { return z / (1.f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = 1.f / (1.f + __expf(-z));
  return s * (1.f + z * (1.f - s));
}

// Z = C + b (in place), H = act(Z) as bf16 (act = SiLU, or identity).
__global__ void k_bias_act(float* __restrict__ Z, const float* __restrict__ bias, int rows, int cols,
                           int act, unsigned short* __restrict__ H) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= (int64_t)rows * cols) return;
  const int c = (int)(e % cols);
  const float z = Z[e] + bias[c];
  Z[e] = z;
  if (H) H[e] = f2bf(act ? silu(z) : z);
}

// dZ = dH * act'(Z); output bf16 (the next products' operand).
__global__ void k_act_grad(const float* __restrict__ dH, const float* __restrict__ Z, int64_t n,
                           unsigned short* __restrict__ dZ) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n) return;
  dZ[e] = f2bf(dH[e] * silu_grad(Z[e]));
}

// db[c] = sum over rows of dZ[r][c] (bf16 values, fixed row order).
__global__ void k_colsum(const unsigned short* __restrict__ dZ, int rows, int cols,
                         float* __restrict__ db) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  float a = 0.f;
  for (int r = 0; r < rows; ++r) a += bf2f(dZ[(int64_t)r * cols + c]);
  db[c] = a;
}

__global__ void k_f32_to_bf16(const float* __restrict__ x, int64_t n, unsigned short* __restrict__ y) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e < n) y[e] = f2bf(x[e]);
}

// Offset features: u = o / |o| (3), 1 / |o|, 1 / |o|^2, u_i u_j (6); zero padded.
__device__ void offset_features(int ox, int oy, int oz, float* f) {
  for (int q = 0; q < kMlpIn; ++q) f[q] = 0.f;
  if (ox == 0 && oy == 0 && oz == 0) return;  // the zero offset is never used
  const float R = sqrtf((float)(ox * ox + oy * oy + oz * oz));
  const float ir = 1.f / R;
  const float u[3] = {ox * ir, oy * ir, oz * ir};
  f[0] = u[0]; f[1] = u[1]; f[2] = u[2];
  f[3] = ir; f[4] = ir * ir;
  f[5] = u[0] * u[0]; f[6] = u[1] * u[1]; f[7] = u[2] * u[2];
  f[8] = u[0] * u[1]; f[9] = u[0] * u[2]; f[10] = u[1] * u[2];
}

__global__ void k_features(const int* __restrict__ off3, int ns, unsigned short* __restrict__ X) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  float f[kMlpIn];
  offset_features(off3[3 * s], off3[3 * s + 1], off3[3 * s + 2], f);
  for (int q = 0; q < kMlpIn; ++q) X[(int64_t)s * kMlpIn + q] = f2bf(f[q]);
}

__device__ __forceinline__ int moment_order(int k) { return k < 9 ? 0 : (k < 36 ? 1 : 2); }
__device__ __forceinline__ int poly_order(int l) { return l == 0 ? 0 : (l < 4 ? 1 : 2); }

// Per sample: prediction, loss term and dY (fp32) of the contraction.
// One workgroup per sample; threads over the 5400 operator entries.
__global__ void k_contract(const float* __restrict__ Y, const int* __restrict__ off3,
                           const float* __restrict__ Ms, const float* __restrict__ Ps,
                           const float* __restrict__ sig, const float* __restrict__ den, int ns,
                           float envScale, float* __restrict__ dY, float* __restrict__ lossPer) {
  const int s = blockIdx.x;
  const int tid = threadIdx.x;
  __shared__ float part[256][6];
  __shared__ float r[6];
  const int ox = off3[3 * s], oy = off3[3 * s + 1], oz = off3[3 * s + 2];
  const float R = sqrtf((float)(ox * ox + oy * oy + oz * oz));
  const float iR = 1.f / R;
  const float* y = Y + (int64_t)s * kMlpOutPad;
  const float* M = Ms + (int64_t)s * kMoments;
  const float* P = Ps + (int64_t)s * kPoly;
  float acc[6] = {0, 0, 0, 0, 0, 0};
  for (int e = tid; e < kChan; e += blockDim.x) {
    const int q = e / kMoments, k = e % kMoments;
    const int comp = q / kPoly, l = q % kPoly;
    const float env = envScale * powf(iR, (float)(2 + moment_order(k) + poly_order(l)));
    acc[comp] += env * y[e] * M[k] * P[l];
  }
  for (int c = 0; c < 6; ++c) part[tid][c] = acc[c];
  __syncthreads();
  if (tid < 6) {
    float a = 0.f;
    for (int t = 0; t < (int)blockDim.x; ++t) a += part[t][tid];
    r[tid] = a - sig[6 * s + tid];
  }
  __syncthreads();
  const float w = 2.f / (den[s] * ns);
  for (int e = tid; e < kMlpOutPad; e += blockDim.x) {
    float g = 0.f;
    if (e < kChan) {
      const int q = e / kMoments, k = e % kMoments;
      const int comp = q / kPoly, l = q % kPoly;
      const float env = envScale * powf(iR, (float)(2 + moment_order(k) + poly_order(l)));
      g = w * r[comp] * env * M[k] * P[l];
    }
    dY[(int64_t)s * kMlpOutPad + e] = g;
  }
  if (tid == 0) {
    float a = 0.f;
    for (int c = 0; c < 6; ++c) a += r[c] * r[c];
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

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float c1,
                       float c2, float eps) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gg = g[e];
  const float mm = b1 * m[e] + (1.f - b1) * gg;
  const float vv = b2 * v[e] + (1.f - b2) * gg * gg;
  m[e] = mm;
  v[e] = vv;
  p[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// Batch b of the training set: sample idx[b] of the resident arrays.
__global__ void k_gather_batch(const int* __restrict__ idx, int nb, const int* __restrict__ off3,
                               const float* __restrict__ M, const float* __restrict__ P,
                               const float* __restrict__ sig, const float* __restrict__ den,
                               int* __restrict__ bOff, float* __restrict__ bM, float* __restrict__ bP,
                               float* __restrict__ bSig, float* __restrict__ bDen) {
  const int b = blockIdx.x;
  const int t = threadIdx.x;
  if (b >= nb) return;
  const int64_t s = idx[b];
  if (t < kMoments) bM[(int64_t)b * kMoments + t] = M[s * kMoments + t];
  if (t < kPoly) bP[(int64_t)b * kPoly + t] = P[s * kPoly + t];
  if (t < 6) bSig[(int64_t)b * 6 + t] = sig[s * 6 + t];
  if (t < 3) bOff[(int64_t)b * 3 + t] = off3[s * 3 + t];
  if (t == 0) bDen[b] = den[s];
}

// Exact stress of source cell c (its segments in cell order) at point x,
// fp64 (the non-singular field of DislocationFEM.jl:142-299).
__global__ void k_sample_stress(const double* __restrict__ seg, int64_t ld,
                                const int* __restrict__ start, const int* __restrict__ order,
                                const int* __restrict__ srcCell, const double* __restrict__ pts,
                                int ns, SigK<double> K, float* __restrict__ sig) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  const int c = srcCell[s];
  const v3<double> x = V(pts[3 * s], pts[3 * s + 1], pts[3 * s + 2]);
  double sg[6] = {0, 0, 0, 0, 0, 0};
  for (int e = start[c]; e < start[c + 1]; ++e) {
    const int i = order[e];
    const v3<double> n1 = V(seg[F_N1X * ld + i], seg[F_N1Y * ld + i], seg[F_N1Z * ld + i]);
    const v3<double> n2 = V(seg[F_N2X * ld + i], seg[F_N2Y * ld + i], seg[F_N2Z * ld + i]);
    const v3<double> b = V(seg[F_BX * ld + i], seg[F_BY * ld + i], seg[F_BZ * ld + i]);
    const v3<double> t = V(seg[F_TX * ld + i], seg[F_TY * ld + i], seg[F_TZ * ld + i]);
    seg_stress<double>(K, x - n1, n2 - n1, t, b, cross(t, b), sg);
  }
  for (int q = 0; q < 6; ++q) sig[6 * s + q] = (float)sg[q];
}

}  // namespace

static unsigned blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

void mlp_gemm(const unsigned short* A, int64_t a_i, int64_t a_k, const unsigned short* B,
              int64_t b_k, int64_t b_j, float* C, int64_t ldc, int M, int N, int K, int accumulate,
              hipStream_t st) {
  dim3 grid(blocks(N, GT), blocks(M, GT));
  hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, st, A, a_i, a_k, B, b_k, b_j, C, ldc, M, N, K,
                     accumulate);
}

void mlp_bias_act(float* Z, const float* b, int rows, int cols, int act, unsigned short* H,
                  hipStream_t st) {
  hipLaunchKernelGGL(k_bias_act, dim3(blocks((int64_t)rows * cols, 256)), dim3(256), 0, st, Z, b,
                     rows, cols, act, H);
}

void mlp_act_grad(const float* dH, const float* Z, int64_t n, unsigned short* dZ, hipStream_t st) {
  hipLaunchKernelGGL(k_act_grad, dim3(blocks(n, 256)), dim3(256), 0, st, dH, Z, n, dZ);
}

void mlp_colsum(const unsigned short* dZ, int rows, int cols, float* db, hipStream_t st) {
  hipLaunchKernelGGL(k_colsum, dim3(blocks(cols, 256)), dim3(256), 0, st, dZ, rows, cols, db);
}

void mlp_to_bf16(const float* x, int64_t n, unsigned short* y, hipStream_t st) {
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(blocks(n, 256)), dim3(256), 0, st, x, n, y);
}

void mlp_features(const int* off3, int ns, unsigned short* X, hipStream_t st) {
  hipLaunchKernelGGL(k_features, dim3(blocks(ns, 256)), dim3(256), 0, st, off3, ns, X);
}

void mlp_contract(const float* Y, const int* off3, const float* M, const float* P, const float* sig,
                  const float* den, int ns, float envScale, float* dY, float* lossPer,
                  hipStream_t st) {
  hipLaunchKernelGGL(k_contract, dim3(ns), dim3(256), 0, st, Y, off3, M, P, sig, den, ns, envScale,
                     dY, lossPer);
}

void mlp_operator_from_y(const float* Y, const int* off3, int no, float envScale, float* W,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_operator_from_y, dim3(blocks((int64_t)no * kChan, 256)), dim3(256), 0, st, Y,
                     off3, no, envScale, W);
}

void mlp_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
              float c1, float c2, float eps, hipStream_t st) {
  hipLaunchKernelGGL(k_adam, dim3(blocks(n, 256)), dim3(256), 0, st, p, g, m, v, n, lr, b1, b2, c1,
                     c2, eps);
}

void gather_batch(const int* idx, int nb, const int* off3, const float* M, const float* P,
                  const float* sig, const float* den, int* bOff, float* bM, float* bP, float* bSig,
                  float* bDen, hipStream_t st) {
  hipLaunchKernelGGL(k_gather_batch, dim3(nb), dim3(128), 0, st, idx, nb, off3, M, P, sig, den, bOff,
                     bM, bP, bSig, bDen);
}

void sample_stress(const double* seg, int64_t ld, const int* start, const int* order,
                   const int* srcCell, const double* pts, int ns, const SigK<double>& K, float* sig,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_sample_stress, dim3(blocks(ns, 128)), dim3(128), 0, st, seg, ld, start, order,
                     srcCell, pts, ns, K, sig);
}

}  // namespace ddd
