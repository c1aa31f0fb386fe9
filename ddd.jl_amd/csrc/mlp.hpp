// Far-field network: sizes and launches (see mlp.hip).
#pragma once
#include "farfield.hpp"

namespace ddd {

constexpr int kMlpIn = 128;        // offset features, zero padded
constexpr int kMonoDeg = 7;        // direction monomials of degree 1..7 (119)
constexpr int kHidden = 512;
constexpr int kMlpOutPad = 5408;   // kChan = 5400 rounded up to a multiple of 32

// Parameter layout (fp32): W1, b1, W2, b2, W3, b3, W4, b4, matrices row-major
// (in x out).
constexpr int64_t kOffW1 = 0;
constexpr int64_t kOffB1 = kOffW1 + (int64_t)kMlpIn * kHidden;
constexpr int64_t kOffW2 = kOffB1 + kHidden;
constexpr int64_t kOffB2 = kOffW2 + (int64_t)kHidden * kHidden;
constexpr int64_t kOffW3 = kOffB2 + kHidden;
constexpr int64_t kOffB3 = kOffW3 + (int64_t)kHidden * kHidden;
constexpr int64_t kOffW4 = kOffB3 + kHidden;
constexpr int64_t kOffB4 = kOffW4 + (int64_t)kHidden * kMlpOutPad;
constexpr int64_t kMlpParams = kOffB4 + kMlpOutPad;

// Epilogue of a layer's product: with a bias, C[i][j] = C[i][j] + bias[j]
// (the pre-activation, kept for the backward pass) and, with H,
// H[i][j] = bf16(SiLU(C[i][j])) (act = 1) or bf16(C[i][j]) (act = 0), both
// with C's row stride.
struct MlpEpilogue {
  const float* bias = nullptr;
  int act = 0;
  unsigned short* H = nullptr;
};
// C[M x N] (+)= A B in BF16 MFMA with FP32 accumulation; A(i, k) = A[i a_i + k a_k],
// B(k, j) = B[k b_k + j b_j], then the epilogue.  work (workFloats floats)
// holds split-K partials when the output has few tiles; pass 0 floats to
// never split.
void mlp_gemm(const unsigned short* A, int64_t a_i, int64_t a_k, const unsigned short* B,
              int64_t b_k, int64_t b_j, float* C, int64_t ldc, int M, int N, int K, int accumulate,
              hipStream_t st, float* work = nullptr, int64_t workFloats = 0,
              MlpEpilogue ep = MlpEpilogue{});
// Column sums of bf16 matrices are formed over chunks of kColRows rows, then
// over the chunks in order; part holds ceil(rows / kColRows) x cols floats.
constexpr int kColRows = 64;
inline int64_t mlp_colsum_part_floats(int rows, int cols) {
  return (int64_t)((rows + kColRows - 1) / kColRows) * cols;
}
// db = column sums of dZ (rows x cols, bf16).
void mlp_colsum(const unsigned short* dZ, int rows, int cols, float* part, float* db,
                hipStream_t st);
// dZ = bf16(dH * SiLU'(Z)) (rows x cols) and db = its column sums.
void mlp_act_grad_colsum(const float* dH, const float* Z, int rows, int cols, unsigned short* dZ,
                         float* part, float* db, hipStream_t st);
void mlp_to_bf16(const float* x, int64_t n, unsigned short* y, hipStream_t st);
// aH: core radius in cell units (features in 1/R_a, R_a^2 = R^2 + aH^2).
void mlp_features(const int* off3, int ns, float aH, unsigned short* X, hipStream_t st);
// Loss terms and dY (bf16, the backward products' operand).
void mlp_contract(const float* Y, const int* off3, const float* M, const float* Lt,
                  const float* den, int ns, float envScale, unsigned short* dY, float* lossPer,
                  hipStream_t st);
void mlp_operator_from_y(const float* Y, const int* off3, int no, float envScale, float* W,
                         hipStream_t st);
// One Adam step of the fp32 parameters p; pb = bf16(p) after the step.
void mlp_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
              float c1, float c2, float eps, unsigned short* pb, hipStream_t st);
// cubic: samples carried into the frame of their canonical offset.
void gather_batch(const int* idx, int nb, const int* off3, const int* src, const float* M, int nc,
                  const float* Lt, const float* den, int cubic, int* bOff, float* bM, float* bL,
                  float* bDen, hipStream_t st);
// W: (2G - 1)^3 x kChan from the canonical operator Wc (C(G + 2, 3) x kChan).
void operator_expand(const float* Wc, int G, float* W, hipStream_t st);

// L2 projection onto the quadratic basis from the 27 tensor Gauss points of a
// cell: node p at (xi[p % 3], xi[(p / 3) % 3], xi[p / 9]) (cell units,
// centred), L_l = sum_p P[l][p] sigma(p).
struct Projection {
  double xi[3];
  double P[kPoly][27];
};
void local_targets(const double* seg, int64_t ld, const GridDev& g, const int* start,
                   const int* order, const int* src, const int* off3, int ns, const SigK<double>& K,
                   const Projection& Pj, float* Lt, hipStream_t st);

}  // namespace ddd
