// The mixed-precision all-pairs kernels (tile kernel and near-parallel
// passes of mode 1, from tiles.hpp).  This translation unit is compiled with
// floating-point contraction and fast FP32 division and square root: the
// mixed path is an approximation held to the FP32 quadrature's accuracy
// (~1e-6 per node against the FP64 path), and its classification of pairs
// uses explicitly rounded arithmetic, so it matches the count and listing
// passes bit for bit.
#include "kernels.hpp"
#include "pair_special.hpp"
#include "tiles.hpp"

namespace ddd {

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
                               unsigned long long* nExact) {
  const int64_t maxGrid = 1ll << 30;
  for (int64_t a = r.t0(); a < r.t1(); a += maxGrid) {
    const int64_t cnt = (r.t1() - a < maxGrid) ? (r.t1() - a) : maxGrid;
    hipLaunchKernelGGL(k_segseg_tiles<1>, dim3((unsigned)cnt), dim3(kTile), 0, st, seg, ld, n, a, K,
                       mr.K, mr.farDist2(), mr.sinMin2(), rowPart, colPart, sp.off, sp.slot, sp.res,
                       sp.total, r.t0(), sp.base, nExact, nullptr, nullptr);
  }
}

}  // namespace ddd
