// Device-resident time integration: the node-level kernels of deriv! and
// integrate! (src/Processing/IntegrationProcessing.jl:13-168).  Built without
// FMA contraction: every operation rounds as in the reference's Julia
// expressions, written out here in the same order.
#include "kernels.hpp"

namespace ddd {

namespace {

// Segment end points from the device coordinates: seg fields n1, n2 of
// segment i from coord[:, nodes[2i]], coord[:, nodes[2i + 1]] (0-based node
// numbers); b and the external flag stay as uploaded.
__global__ void k_gather_coords(double* __restrict__ seg, int64_t ld, int64_t n,
                                const int64_t* __restrict__ nodes,
                                const double* __restrict__ coord) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* a = coord + 3 * nodes[2 * i];
  const double* b = coord + 3 * nodes[2 * i + 1];
  seg[F_N1X * ld + i] = a[0];
  seg[F_N1Y * ld + i] = a[1];
  seg[F_N1Z * ld + i] = a[2];
  seg[F_N2X * ld + i] = b[0];
  seg[F_N2Y * ld + i] = b[1];
  seg[F_N2Z * ld + i] = b[2];
}

// Forward Euler trial (IntegrationProcessing.jl:123): coord = init + vel dt.
__global__ void k_euler(double* __restrict__ coord, const double* __restrict__ init,
                        const double* __restrict__ vel, double dt, int64_t m) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= m) return;
  coord[e] = init[e] + vel[e] * dt;
}

constexpr int kRedBlock = 256;
constexpr int kRedGrid = 1024;

// Largest |distance| and |distance - (vel + vel0) / 2 dt| over all entries
// (IntegrationProcessing.jl:131-136), per block.  NaN propagates, as
// Julia's maximum does: |NaN| has the largest bit pattern of all
// non-negative doubles, and the maximum is taken on bit patterns.
__global__ void __launch_bounds__(kRedBlock)
k_trapezoid_max(const double* __restrict__ coord, const double* __restrict__ init,
                const double* __restrict__ vel, const double* __restrict__ vel0, double dt,
                int64_t m, unsigned long long* __restrict__ part) {
  __shared__ unsigned long long sd[kRedBlock], se[kRedBlock];
  unsigned long long md = 0, me = 0;
  for (int64_t e = blockIdx.x * (int64_t)kRedBlock + threadIdx.x; e < m;
       e += (int64_t)gridDim.x * kRedBlock) {
    const double dist = coord[e] - init[e];
    const double err = dist - (vel[e] + vel0[e]) / 2.0 * dt;
    const unsigned long long bd = (unsigned long long)__double_as_longlong(fabs(dist));
    const unsigned long long be = (unsigned long long)__double_as_longlong(fabs(err));
    md = bd > md ? bd : md;
    me = be > me ? be : me;
  }
  sd[threadIdx.x] = md;
  se[threadIdx.x] = me;
  __syncthreads();
  for (int s = kRedBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sd[threadIdx.x] = sd[threadIdx.x + s] > sd[threadIdx.x] ? sd[threadIdx.x + s] : sd[threadIdx.x];
      se[threadIdx.x] = se[threadIdx.x + s] > se[threadIdx.x] ? se[threadIdx.x + s] : se[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sd[0];
    part[2 * blockIdx.x + 1] = se[0];
  }
}

__global__ void __launch_bounds__(kRedBlock)
k_trapezoid_max_final(const unsigned long long* __restrict__ part, int nPart,
                      double* __restrict__ out) {
  __shared__ unsigned long long sd[kRedBlock], se[kRedBlock];
  unsigned long long md = 0, me = 0;
  for (int b = threadIdx.x; b < nPart; b += kRedBlock) {
    md = part[2 * b] > md ? part[2 * b] : md;
    me = part[2 * b + 1] > me ? part[2 * b + 1] : me;
  }
  sd[threadIdx.x] = md;
  se[threadIdx.x] = me;
  __syncthreads();
  for (int s = kRedBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sd[threadIdx.x] = sd[threadIdx.x + s] > sd[threadIdx.x] ? sd[threadIdx.x + s] : sd[threadIdx.x];
      se[threadIdx.x] = se[threadIdx.x + s] > se[threadIdx.x] ? se[threadIdx.x + s] : se[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = __longlong_as_double((long long)sd[0]);
    out[1] = __longlong_as_double((long long)se[0]);
  }
}

}  // namespace

void launch_gather_coords(double* seg, int64_t ld, int64_t n, const int64_t* nodes,
                          const double* coord, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gather_coords, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, seg, ld,
                     n, nodes, coord);
}

void launch_euler(double* coord, const double* init, const double* vel, double dt, int64_t m,
                  hipStream_t st) {
  if (m <= 0) return;
  hipLaunchKernelGGL(k_euler, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, coord, init,
                     vel, dt, m);
}

int trapezoid_parts() { return kRedGrid; }

void launch_trapezoid_max(const double* coord, const double* init, const double* vel,
                          const double* vel0, double dt, int64_t m, unsigned long long* part,
                          double* out, hipStream_t st) {
  const int64_t want = (m + kRedBlock - 1) / kRedBlock;
  const int grid = (int)(want < kRedGrid ? (want > 0 ? want : 1) : kRedGrid);
  hipLaunchKernelGGL(k_trapezoid_max, dim3(grid), dim3(kRedBlock), 0, st, coord, init, vel, vel0, dt,
                     m, part);
  hipLaunchKernelGGL(k_trapezoid_max_final, dim3(1), dim3(kRedBlock), 0, st, part, grid, out);
}

}  // namespace ddd
