// Far-field network: sizes and launches (see mlp.hip).
#pragma once
#include "farfield.hpp"

namespace ddd {

constexpr int kMlpIn = 32;         // offset features, zero padded
constexpr int kHidden = 256;
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

void mlp_gemm(const unsigned short* A, int64_t a_i, int64_t a_k, const unsigned short* B,
              int64_t b_k, int64_t b_j, float* C, int64_t ldc, int M, int N, int K, int accumulate,
              hipStream_t st);
void mlp_bias_act(float* Z, const float* b, int rows, int cols, int act, unsigned short* H,
                  hipStream_t st);
void mlp_act_grad(const float* dH, const float* Z, int64_t n, unsigned short* dZ, hipStream_t st);
void mlp_colsum(const unsigned short* dZ, int rows, int cols, float* db, hipStream_t st);
void mlp_to_bf16(const float* x, int64_t n, unsigned short* y, hipStream_t st);
void mlp_features(const int* off3, int ns, unsigned short* X, hipStream_t st);
void mlp_contract(const float* Y, const int* off3, const float* M, const float* P, const float* sig,
                  const float* den, int ns, float envScale, float* dY, float* lossPer,
                  hipStream_t st);
void mlp_operator_from_y(const float* Y, const int* off3, int no, float envScale, float* W,
                         hipStream_t st);
void mlp_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
              float c1, float c2, float eps, hipStream_t st);
void gather_batch(const int* idx, int nb, const int* off3, const float* M, const float* P,
                  const float* sig, const float* den, int* bOff, float* bM, float* bP, float* bSig,
                  float* bDen, hipStream_t st);
void sample_stress(const double* seg, int64_t ld, const int* start, const int* order,
                   const int* srcCell, const double* pts, int ns, const SigK<double>& K, float* sig,
                   hipStream_t st);

}  // namespace ddd
