// Hybrid near/far field: cell grid, moments, far-field operator, launches.
#pragma once
#include "kernels.hpp"
#include "pair_special.hpp"

namespace ddd {

constexpr int kMoments = 90;  // cell features: orders 0, 1, 2 of b (x) t L
constexpr int kPoly = 10;     // local basis: 1, x, y, z, x^2, y^2, z^2, xy, xz, yz
constexpr int kLocal = 6 * kPoly;          // stress components x basis
constexpr int kChan = kLocal * kMoments;   // entries of one W(o)

struct GridDev {
  double lo[3];
  double h;
  int G;
};

int fft_size_for(int G);
void launch_cells(const double* seg, int64_t ld, int n, const GridDev& g, int* cellOf,
                  unsigned* counts, unsigned* cursor, int* start, int* tmp, int* order,
                  hipStream_t st);
void launch_moments(const double* seg, int64_t ld, const GridDev& g, const int* start,
                    const int* order, float* M, hipStream_t st);
void launch_far_field(const float* M, const GridDev& g, const float2* What, const float2* tw,
                      float2* Mc, float2* Lc, float* L, hipStream_t st);
void launch_operator_spectrum(const float* W, const GridDev& g, int rn, const float2* tw,
                              float2* scratch, float2* What, hipStream_t st);
void launch_far_apply(const double* seg, int64_t ld, int n, const GridDev& g, const int* cellOf,
                      const float* L, double* out, hipStream_t st);
void launch_near_rows(const double* seg, int64_t ld, const GridDev& g, int rn, const int* start,
                      const int* order, const SegK<double>& K, const MixedRule& mr, double* out,
                      unsigned* flags, unsigned long long* counters, hipStream_t st);

}  // namespace ddd
