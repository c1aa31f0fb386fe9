// Hybrid force path (no reference counterpart): exact near field on a cell
// grid plus a far field from per-cell dislocation moments and a learned
// linear operator applied as a 3-D convolution through FFTs.
//
// Cells.  Segments are binned by midpoint into a G^3 grid of edge h starting
// at lo (clamped at the borders) and ordered by (cell, segment index).
// Near pairs are those whose cells are within Chebyshev distance rn; they
// are evaluated exactly (FP64 analytic, or the FP32 far-pair quadrature for
// well-separated non-parallel pairs).  All other pairs form the far field.
//
// Moments.  For cell c with centre x_c, with r = (mid - x_c)/h and
// l = t L / h (unit tangent times length in cell units), the 90 features are
//   order 0:  b_i l_j                                   (9)
//   order 1:  b_i l_j r_k                               (27)
//   order 2:  b_i l_j (r_k r_l + l_k l_l / 12), k <= l  (54)
// summed over the cell's segments (the second-order term includes the
// segment's own extent).
//
// Far field.  The stress in target cell t is represented as a quadratic
// polynomial in rho = (x - x_t)/h with 6 x 10 coefficients
//   L_t = sum over far source cells s of W(s - t) M_s,   W(o) in R^{60 x 90},
// a correlation over the grid evaluated by FFT on a (2G)^3 periodic grid,
// exact for G^3 data with zero padding.  W comes from the far-field model.
// The nodal force on a segment is the two-point Gauss-Legendre quadrature of
// (sigma b) x t against its shape functions, as in far_side.
#include "kernels.hpp"
#include "farfield.hpp"

#include <algorithm>
#include <vector>

namespace ddd {

namespace {

__device__ __forceinline__ int cell_of(const GridDev& g, double x, double y, double z) {
  int cx = (int)floor((x - g.lo[0]) / g.h), cy = (int)floor((y - g.lo[1]) / g.h),
      cz = (int)floor((z - g.lo[2]) / g.h);
  cx = cx < 0 ? 0 : (cx >= g.G ? g.G - 1 : cx);
  cy = cy < 0 ? 0 : (cy >= g.G ? g.G - 1 : cy);
  cz = cz < 0 ? 0 : (cz >= g.G ? g.G - 1 : cz);
  return cx + g.G * (cy + g.G * cz);
}

__global__ void k_cell_assign(const double* __restrict__ seg, int64_t ld, int n, GridDev g,
                              int* __restrict__ cellOf, unsigned* __restrict__ counts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (seg[F_EXT * ld + i] != 0.0) {
    cellOf[i] = -1;
    return;
  }
  const double mx = 0.5 * (seg[F_N1X * ld + i] + seg[F_N2X * ld + i]);
  const double my = 0.5 * (seg[F_N1Y * ld + i] + seg[F_N2Y * ld + i]);
  const double mz = 0.5 * (seg[F_N1Z * ld + i] + seg[F_N2Z * ld + i]);
  const int c = cell_of(g, mx, my, mz);
  cellOf[i] = c;
  atomicAdd(&counts[c], 1u);
}

// Exclusive scan of nc counts into start[0..nc] by one workgroup of 1,024:
// tiles of 1,024 counts, read coalesced, scanned within each wavefront by
// shuffles and across the 16 wavefronts through LDS, with a running carry.
__global__ __launch_bounds__(1024) void k_cell_scan(const unsigned* __restrict__ counts, int nc,
                                                    int* __restrict__ start) {
  __shared__ int wsum[16];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  int carry = 0;
  for (int base = 0; base < nc; base += 1024) {
    const int c = base + tid;
    const int v = c < nc ? (int)counts[c] : 0;
    int x = v;  // inclusive within the wavefront
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int y = __shfl_up(x, d);
      if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    if (w == 0) {  // inclusive over the wavefronts' totals
      int t = lane < 16 ? wsum[lane] : 0;
#pragma unroll
      for (int d = 1; d < 16; d <<= 1) {
        const int y = __shfl_up(t, d);
        if (lane >= d) t += y;
      }
      if (lane < 16) wsum[lane] = t;
    }
    __syncthreads();
    if (c < nc) start[c] = carry + (w ? wsum[w - 1] : 0) + x - v;
    carry += wsum[15];
    __syncthreads();  // wsum is rewritten by the next tile
  }
  if (tid == 0) start[nc] = carry;
}

__global__ void k_cell_scatter(const int* __restrict__ cellOf, int n, const int* __restrict__ start,
                               unsigned* __restrict__ cursor, int* __restrict__ tmp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || cellOf[i] < 0) return;
  const int c = cellOf[i];
  tmp[start[c] + (int)atomicAdd(&cursor[c], 1u)] = i;
}

// Order each cell's segments by index (the scatter's order is arbitrary):
// the rank of an entry is the number of smaller indices in its cell.
__global__ void k_cell_rank(const int* __restrict__ start, const int* __restrict__ tmp,
                            int* __restrict__ order) {
  const int c = blockIdx.x;
  const int s0 = start[c], s1 = start[c + 1];
  for (int e = s0 + threadIdx.x; e < s1; e += blockDim.x) {
    const int v = tmp[e];
    int r = 0;
    for (int f = s0; f < s1; ++f) r += tmp[f] < v;
    order[s0 + r] = v;
  }
}

// Feature f of the cell moments (see the file comment) for one segment.
__device__ __forceinline__ double moment_feature(int f, const double b[3], const double l[3],
                                                 const double r[3]) {
  if (f < 9) return b[f / 3] * l[f % 3];
  if (f < 36) {
    const int g = f - 9;
    return b[g / 9] * l[(g / 3) % 3] * r[g % 3];
  }
  const int g = f - 36;
  const int KK[6] = {0, 1, 2, 0, 0, 1}, LL[6] = {0, 1, 2, 1, 2, 2};
  const int k = KK[g % 6], q = LL[g % 6];
  return b[g / 18] * l[(g / 6) % 3] * (r[k] * r[q] + l[k] * l[q] / 12.0);
}

// One workgroup per cell, one thread per feature; segments in cell order.
__global__ void k_cell_moments(const double* __restrict__ seg, int64_t ld, GridDev g,
                               const int* __restrict__ start, const int* __restrict__ order,
                               float* __restrict__ M) {
  const int c = blockIdx.x;
  const int f = threadIdx.x;
  if (f >= kMoments) return;
  const int cx = c % g.G, cy = (c / g.G) % g.G, cz = c / (g.G * g.G);
  const double xc = g.lo[0] + (cx + 0.5) * g.h, yc = g.lo[1] + (cy + 0.5) * g.h,
               zc = g.lo[2] + (cz + 0.5) * g.h;
  const double ih = 1.0 / g.h;
  double acc = 0.0;
  for (int e = start[c]; e < start[c + 1]; ++e) {
    const int i = order[e];
    const double n1[3] = {seg[F_N1X * ld + i], seg[F_N1Y * ld + i], seg[F_N1Z * ld + i]};
    const double n2[3] = {seg[F_N2X * ld + i], seg[F_N2Y * ld + i], seg[F_N2Z * ld + i]};
    const double b[3] = {seg[F_BX * ld + i], seg[F_BY * ld + i], seg[F_BZ * ld + i]};
    const double l[3] = {(n2[0] - n1[0]) * ih, (n2[1] - n1[1]) * ih, (n2[2] - n1[2]) * ih};
    const double r[3] = {(0.5 * (n1[0] + n2[0]) - xc) * ih, (0.5 * (n1[1] + n2[1]) - yc) * ih,
                         (0.5 * (n1[2] + n2[2]) - zc) * ih};
    acc += moment_feature(f, b, l, r);
  }
  M[(int64_t)f * g.G * g.G * g.G + c] = (float)acc;
}

// ---- 3-D FFT on an NP^3 periodic grid, one line per thread ----------------

// Workgroup size of the line transforms.  Stated in the kernel's launch
// bounds: without them the compiler budgets for 1,024-thread workgroups (128
// registers a lane) and a 64-point line, held in 128 registers, spills.
constexpr int kFftBlock = 128;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// In-register radix-2 transform of one line whose element k sits at
// p[k stride] (loaded in bit-reversed order, written back in order).
template <int NP>
__device__ __forceinline__ void fft_line(float2* p, int64_t stride, int inverse,
                                         const float2* __restrict__ tw) {
  constexpr int LOG = NP == 32 ? 5 : (NP == 64 ? 6 : (NP == 16 ? 4 : 3));
  float2 v[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    const int r = __brev(k) >> (32 - LOG);
    v[r] = p[k * stride];
  }
#pragma unroll
  for (int len = 2; len <= NP; len <<= 1) {
    const int half = len >> 1, step = NP / len;
#pragma unroll
    for (int i = 0; i < NP; i += len) {
#pragma unroll
      for (int j = 0; j < half; ++j) {
        float2 w = tw[j * step];
        if (inverse) w.y = -w.y;
        const float2 u = v[i + j], t = cmul(w, v[i + j + half]);
        v[i + j] = make_float2(u.x + t.x, u.y + t.y);
        v[i + j + half] = make_float2(u.x - t.x, u.y - t.y);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < NP; ++k) p[k * stride] = v[k];
}

// Lines along y (axis 1) or z (axis 2), one a thread: consecutive threads
// take consecutive x, so each of a line's loads is coalesced across lanes.
template <int NP>
__global__ __launch_bounds__(kFftBlock) void k_fft_lines(float2* __restrict__ data, int64_t nChan,
                                                         int axis, int inverse,
                                                         const float2* __restrict__ tw) {
  const int64_t line = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)NP * NP;
  if (line >= nChan * per) return;
  const int64_t ch = line / per;
  const int rem = (int)(line % per);
  const int a = rem % NP, b = rem / NP;
  // element (x, y, z) of channel ch at ch NP^3 + x + NP (y + NP z)
  const int64_t base = axis == 1 ? a + per * b : a + (int64_t)NP * b;
  const int64_t stride = axis == 1 ? NP : per;
  fft_line<NP>(data + ch * per * NP + base, stride, inverse, tw);
}

// Lines along x (axis 0): a workgroup of NP threads takes NP consecutive
// lines, one contiguous NP x NP block, loaded and stored coalesced through
// LDS (rows padded by one element against bank conflicts); each thread
// transforms one line from LDS as k_fft_lines does.
template <int NP>
__global__ __launch_bounds__(NP) void k_fft_rows(float2* __restrict__ data, int inverse,
                                                 const float2* __restrict__ tw) {
  __shared__ float2 sm[NP][NP + 1];
  const int t = threadIdx.x;
  float2* blk = data + (int64_t)blockIdx.x * NP * NP;
#pragma unroll 8
  for (int i = t; i < NP * NP; i += NP) sm[i / NP][i % NP] = blk[i];
  __syncthreads();
  fft_line<NP>(&sm[t][0], 1, inverse, tw);
  __syncthreads();
#pragma unroll 8
  for (int i = t; i < NP * NP; i += NP) blk[i] = sm[i / NP][i % NP];
}

// Moments into the zero-padded complex grid, two real channels per complex
// channel: channel p holds M_{2p} + i M_{2p+1}.
__global__ void k_pad_moments(const float* __restrict__ M, int G, int NP, float2* __restrict__ out) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t vol = (int64_t)NP * NP * NP;
  if (e >= (int64_t)kMomPairs * vol) return;
  const int p = (int)(e / vol);
  const int r = (int)(e % vol);
  const int x = r % NP, y = (r / NP) % NP, z = r / (NP * NP);
  float2 v = make_float2(0.f, 0.f);
  if (x < G && y < G && z < G) {
    const int64_t c = x + G * (y + G * z), nc = (int64_t)G * G * G;
    v = make_float2(M[(2 * p) * nc + c], M[(2 * p + 1) * nc + c]);
  }
  out[e] = v;
}

// Frequency index of -f on the periodic grid.
__device__ __forceinline__ int64_t neg_freq(int fx, int fy, int fz, int NP) {
  const int gx = (NP - fx) & (NP - 1), gy = (NP - fy) & (NP - 1), gz = (NP - fz) & (NP - 1);
  return gx + (int64_t)NP * (gy + (int64_t)NP * gz);
}

// The spectra of the two real channels packed in Z: with a = Z(f) and
// b = conj(Z(-f)), X_even(f) = (a + b) / 2 and X_odd(f) = (a - b) / (2i).
__device__ __forceinline__ float2 unpack_even(float2 a, float2 b) {
  return make_float2(0.5f * (a.x + b.x), 0.5f * (a.y + b.y));
}
__device__ __forceinline__ float2 unpack_odd(float2 a, float2 b) {
  return make_float2(0.5f * (a.y - b.y), -0.5f * (a.x - b.x));
}

// L(f) = W(f) M(f) at the half spectrum f_z in [0, NP/2] (real fields: the
// rest is conjugate), one workgroup per frequency.  The local expansions are
// written packed two per complex channel at f and, off the self-conjugate
// planes f_z = 0 and NP/2, at -f, so every frequency is written once.
__global__ void k_spectral_apply(const float2* __restrict__ What, const float2* __restrict__ Z,
                                 int NP, float2* __restrict__ Y) {
  const int64_t fh = blockIdx.x;  // half-spectrum index: fx + NP (fy + NP fz)
  const int fx = (int)(fh % NP), fy = (int)((fh / NP) % NP), fz = (int)(fh / ((int64_t)NP * NP));
  const int64_t vol = (int64_t)NP * NP * NP;
  const int64_t f = fx + (int64_t)NP * (fy + (int64_t)NP * fz);
  const int64_t g = neg_freq(fx, fy, fz, NP);
  const int q = threadIdx.x;
  __shared__ float2 m[kMoments];
  __shared__ float2 l[kLocal];
  for (int p = threadIdx.x; p < kMomPairs; p += blockDim.x) {
    const float2 a = Z[(int64_t)p * vol + f], zb = Z[(int64_t)p * vol + g];
    const float2 b = make_float2(zb.x, -zb.y);
    m[2 * p] = unpack_even(a, b);
    m[2 * p + 1] = unpack_odd(a, b);
  }
  __syncthreads();
  if (q < kLocal) {
    const float2* w = What + fh * kChan + q;
    float ax = 0.f, ay = 0.f;
    for (int k = 0; k < kMoments; ++k) {
      const float2 a = w[k * kLocal], b = m[k];
      ax += a.x * b.x - a.y * b.y;
      ay += a.x * b.y + a.y * b.x;
    }
    l[q] = make_float2(ax, ay);
  }
  __syncthreads();
  if (q < kLocPairs) {
    const float2 a = l[2 * q], b = l[2 * q + 1];
    Y[(int64_t)q * vol + f] = make_float2(a.x - b.y, a.y + b.x);
    if (fz != 0 && 2 * fz != NP) Y[(int64_t)q * vol + g] = make_float2(a.x + b.y, b.x - a.y);
  }
}

// The inverse transform at the G^3 cells, scaled by 1/NP^3: local channel
// q is the real (q even) or imaginary (q odd) part of packed channel q / 2.
__global__ void k_unpad_local(const float2* __restrict__ Lc, int G, int NP, float* __restrict__ L) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t nc = (int64_t)G * G * G;
  if (e >= (int64_t)kLocal * nc) return;
  const int q = (int)(e / nc);
  const int c = (int)(e % nc);
  const int x = c % G, y = (c / G) % G, z = c / (G * G);
  const float scale = 1.0f / ((float)NP * NP * NP);
  const float2 v = Lc[(int64_t)(q / 2) * NP * NP * NP + x + NP * (y + NP * z)];
  L[e] = (q % 2 == 0 ? v.x : v.y) * scale;
}

// W(o) (offsets o in [-(G-1), G-1]^3, channel-major [o][60][90]) into the
// periodic grid at index -o mod NP, zero inside the near stencil, two real
// channels per complex channel (p holds channels 2p + i (2p + 1)).
__global__ void k_place_operator(const float* __restrict__ W, int G, int NP, int rn,
                                 float2* __restrict__ out) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t vol = (int64_t)NP * NP * NP;
  if (e >= (int64_t)kChanPairs * vol) return;
  const int p = (int)(e / vol);
  const int r = (int)(e % vol);
  const int x = r % NP, y = (r / NP) % NP, z = r / (NP * NP);
  // index x holds offset o = -x (mod NP), folded into (-NP/2, NP/2]
  auto off = [&](int i) { int o = -i; if (o <= -NP / 2) o += NP; return o; };
  const int ox = off(x), oy = off(y), oz = off(z);
  float2 v = make_float2(0.f, 0.f);
  const int S = 2 * G - 1;
  if (abs(ox) <= G - 1 && abs(oy) <= G - 1 && abs(oz) <= G - 1 &&
      max(abs(ox), max(abs(oy), abs(oz))) > rn) {
    const int64_t oi = (ox + G - 1) + (int64_t)S * ((oy + G - 1) + (int64_t)S * (oz + G - 1));
    v = make_float2(W[oi * kChan + 2 * p], W[oi * kChan + 2 * p + 1]);
  }
  out[e] = v;
}

// Packed channel-major spectrum -> [half frequency][moment k][local q] of
// the real channels ch = q * kMoments + k (f_z in [0, NP/2]): per frequency
// the operator is stored column by column, so the 60 lanes of
// k_spectral_apply read one column, 480 contiguous bytes, per step.
__global__ void k_transpose_spectrum(const float2* __restrict__ in, int NP,
                                     float2* __restrict__ out) {
  const int64_t volH = (int64_t)NP * NP * (NP / 2 + 1);
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= (int64_t)kChan * volH) return;
  const int64_t fh = e / kChan;
  const int s = (int)(e % kChan);
  const int ch = (s % kLocal) * kMoments + s / kLocal;
  const int fx = (int)(fh % NP), fy = (int)((fh / NP) % NP), fz = (int)(fh / ((int64_t)NP * NP));
  const int64_t vol = (int64_t)NP * NP * NP;
  const int64_t f = fx + (int64_t)NP * (fy + (int64_t)NP * fz);
  const int64_t g = neg_freq(fx, fy, fz, NP);
  const float2* z = in + (int64_t)(ch / 2) * vol;
  const float2 a = z[f], zb = z[g];
  const float2 b = make_float2(zb.x, -zb.y);
  out[e] = (ch % 2 == 0) ? unpack_even(a, b) : unpack_odd(a, b);
}

__device__ __forceinline__ void local_basis(double x, double y, double z, double P[kPoly]) {
  P[0] = 1.0; P[1] = x; P[2] = y; P[3] = z;
  P[4] = x * x; P[5] = y * y; P[6] = z * z; P[7] = x * y; P[8] = x * z; P[9] = y * z;
}

// Far-field force on every live segment from its cell's local expansion.
__global__ void k_far_apply(const double* __restrict__ seg, int64_t ld, int n, GridDev g,
                            const int* __restrict__ cellOf, const float* __restrict__ L,
                            double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cellOf[i];
  if (c < 0) return;
  const int64_t nc = (int64_t)g.G * g.G * g.G;
  const int cx = c % g.G, cy = (c / g.G) % g.G, cz = c / (g.G * g.G);
  const double xc = g.lo[0] + (cx + 0.5) * g.h, yc = g.lo[1] + (cy + 0.5) * g.h,
               zc = g.lo[2] + (cz + 0.5) * g.h;
  const Seg<double> A = seg_from(seg, ld, i);
  const double gq[2] = {0.21132486540518711775, 0.78867513459481288225};
  v3<double> f[2];
  for (int q = 0; q < 2; ++q) {
    const double x = A.n1.x + gq[q] * (A.n2.x - A.n1.x);
    const double y = A.n1.y + gq[q] * (A.n2.y - A.n1.y);
    const double z = A.n1.z + gq[q] * (A.n2.z - A.n1.z);
    double P[kPoly];
    local_basis((x - xc) / g.h, (y - yc) / g.h, (z - zc) / g.h, P);
    double s[6];
    for (int comp = 0; comp < 6; ++comp) {
      double a = 0.0;
      for (int l = 0; l < kPoly; ++l) a += (double)L[(int64_t)(comp * kPoly + l) * nc + c] * P[l];
      s[comp] = a;
    }
    const v3<double> sb = V(s[0] * A.b.x + s[3] * A.b.y + s[4] * A.b.z,
                            s[3] * A.b.x + s[1] * A.b.y + s[5] * A.b.z,
                            s[4] * A.b.x + s[5] * A.b.y + s[2] * A.b.z);
    f[q] = cross(sb, A.t) * (0.5 / A.tN);
  }
  const v3<double> F1 = f[0] * gq[1] + f[1] * gq[0];
  const v3<double> F2 = f[0] * gq[0] + f[1] * gq[1];
  double* o = out + 6 * (int64_t)i;
  o[0] += F1.x; o[1] += F1.y; o[2] += F1.z;
  o[3] += F2.x; o[4] += F2.y; o[5] += F2.z;
}

// ---- Near field: cell-sorted blocks and their tile list ------------------
//
// Binned segments are copied in Morton order of their cells (ties by
// segment index) into a sorted SoA copy, cut into blocks of 64.  A block pair
// (I <= J) is a near tile when the cell boxes of the two blocks are within
// Chebyshev distance rn; the tile kernel then takes only the pairs whose own
// cells are.  Lists are built per row block with ballots, so their order is
// fixed and the reduction below is deterministic.

__device__ __forceinline__ int key_x(int k) { return k & 1023; }
__device__ __forceinline__ int key_y(int k) { return (k >> 10) & 1023; }
__device__ __forceinline__ int key_z(int k) { return k >> 20; }

__global__ void k_morton_counts(const unsigned* __restrict__ counts, const int* __restrict__ rankCell,
                                int nc, unsigned* __restrict__ mcount) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < nc) mcount[r] = counts[rankCell[r]];
}

__global__ void k_sort_gather(const double* __restrict__ seg, int64_t ld, int n,
                              const int* __restrict__ cellOf, const int* __restrict__ start,
                              const int* __restrict__ order, const int* __restrict__ mrank,
                              const int* __restrict__ mstart, int G, int64_t ldS,
                              double* __restrict__ segS, int* __restrict__ key,
                              int* __restrict__ perm) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n || e >= start[G * G * G]) return;
  const int i = order[e];
  const int c = cellOf[i];
  const int p = mstart[mrank[c]] + (e - start[c]);
  perm[p] = i;
#pragma unroll
  for (int f = 0; f < F_COUNT; ++f) segS[f * ldS + p] = seg[f * ld + i];
  key[p] = (c % G) | (((c / G) % G) << 10) | ((c / (G * G)) << 20);
}

// Positions past the binned segments: inert unit segments marked external.
__global__ void k_sort_pad(const int* __restrict__ start, int nc, int64_t ldS,
                           double* __restrict__ segS, int* __restrict__ key, int* __restrict__ perm) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= ldS || p < start[nc]) return;
  for (int f = 0; f < F_COUNT; ++f) segS[f * ldS + p] = 0.0;
  segS[F_N2X * ldS + p] = 1.0;
  segS[F_TX * ldS + p] = 1.0;
  segS[F_TN * ldS + p] = 1.0;
  segS[F_EXT * ldS + p] = 1.0;
  key[p] = 0;
  perm[p] = -1;
}

// Cell box of each block (lo > hi for an empty block).
__global__ void k_block_boxes(const int* __restrict__ key, const int* __restrict__ start, int nc,
                              int nbS, int* __restrict__ box) {
  const int B = blockIdx.x * blockDim.x + threadIdx.x;
  if (B >= nbS) return;
  const int nbin = start[nc];
  int lo[3] = {1 << 20, 1 << 20, 1 << 20}, hi[3] = {-1, -1, -1};
  const int p1 = min(nbin, 64 * (B + 1));
  for (int p = 64 * B; p < p1; ++p) {
    const int k = key[p];
    const int x[3] = {key_x(k), key_y(k), key_z(k)};
    for (int a = 0; a < 3; ++a) {
      lo[a] = min(lo[a], x[a]);
      hi[a] = max(hi[a], x[a]);
    }
  }
  for (int a = 0; a < 3; ++a) {
    box[6 * B + a] = lo[a];
    box[6 * B + 3 + a] = hi[a];
  }
}

__device__ __forceinline__ bool blocks_near(const int* __restrict__ box, int I, int J, int rn) {
  bool ok = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int loI = box[6 * I + a], hiI = box[6 * I + 3 + a];
    const int loJ = box[6 * J + a], hiJ = box[6 * J + 3 + a];
    ok = ok && loI <= hiI && loJ <= hiJ && (loJ - hiI) <= rn && (loI - hiJ) <= rn;
  }
  return ok;
}

// One wavefront per row block I: number of near J >= I.
__global__ __launch_bounds__(64) void k_near_count_rows(const int* __restrict__ box, int nbS, int rn,
                                                       unsigned* __restrict__ rowCount) {
  const int I = blockIdx.x, lane = threadIdx.x;
  unsigned cnt = 0;
  for (int J0 = I; J0 < nbS; J0 += 64) {
    const int J = J0 + lane;
    cnt += (unsigned)__popcll(__ballot(J < nbS && blocks_near(box, I, J, rn)));
  }
  if (lane == 0) rowCount[I] = cnt;
}

// Tile list in row-block order; tileIdx[I][J] = tile number or -1 (J >= I).
__global__ __launch_bounds__(64) void k_near_fill_rows(const int* __restrict__ box, int nbS, int rn,
                                                      const int* __restrict__ rowStart,
                                                      int2* __restrict__ list,
                                                      int* __restrict__ tileIdx) {
  const int I = blockIdx.x, lane = threadIdx.x;
  const uint64_t ltMask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int pos = rowStart[I];
  for (int J0 = I; J0 < nbS; J0 += 64) {
    const int J = J0 + lane;
    const bool nr = J < nbS && blocks_near(box, I, J, rn);
    const uint64_t m = __ballot(nr);
    if (J < nbS) {
      int t = -1;
      if (nr) {
        t = pos + __popcll(m & ltMask);
        list[t] = make_int2(I, J);
      }
      tileIdx[(int64_t)I * nbS + J] = t;
    }
    pos += __popcll(m);
  }
}

// One wavefront per column block J: number of near I <= J ...
__global__ __launch_bounds__(64) void k_near_count_cols(const int* __restrict__ tileIdx, int nbS,
                                                       unsigned* __restrict__ colCount) {
  const int J = blockIdx.x, lane = threadIdx.x;
  unsigned cnt = 0;
  for (int I0 = 0; I0 <= J; I0 += 64) {
    const int I = I0 + lane;
    cnt += (unsigned)__popcll(__ballot(I <= J && tileIdx[(int64_t)I * nbS + J] >= 0));
  }
  if (lane == 0) colCount[J] = cnt;
}

// ... and their tile numbers in increasing I.
__global__ __launch_bounds__(64) void k_near_fill_cols(const int* __restrict__ tileIdx, int nbS,
                                                      const int* __restrict__ colStart,
                                                      int* __restrict__ colList) {
  const int J = blockIdx.x, lane = threadIdx.x;
  const uint64_t ltMask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int pos = colStart[J];
  for (int I0 = 0; I0 <= J; I0 += 64) {
    const int I = I0 + lane;
    const int t = I <= J ? tileIdx[(int64_t)I * nbS + J] : -1;
    const uint64_t m = __ballot(t >= 0);
    if (t >= 0) colList[pos + __popcll(m & ltMask)] = t;
    pos += __popcll(m);
  }
}

// Sum of a sorted segment's tile partials (row role, then column role, in
// list order) into out at its original index.
__global__ void k_near_reduce(const double* __restrict__ rowPart, const double* __restrict__ colPart,
                              const int* __restrict__ rowStart, const int* __restrict__ colStart,
                              const int* __restrict__ colList, const int* __restrict__ perm,
                              const int* __restrict__ start, int nc, double* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= start[nc]) return;
  const int B = p / kTile, l = p % kTile;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  // unrolled so that several tiles' loads are in flight (the sums stay in
  // list order)
#pragma unroll 8
  for (int t = rowStart[B]; t < rowStart[B + 1]; ++t) {
    const double* r = rowPart + (int64_t)t * 6 * kTile + l;
#pragma unroll
    for (int c = 0; c < 6; ++c) acc[c] += r[c * kTile];
  }
#pragma unroll 8
  for (int q = colStart[B]; q < colStart[B + 1]; ++q) {
    const double* r = colPart + (int64_t)colList[q] * 6 * kTile + l;
#pragma unroll
    for (int c = 0; c < 6; ++c) acc[c] += r[c * kTile];
  }
  double* o = out + 6 * (int64_t)perm[p];
#pragma unroll
  for (int c = 0; c < 6; ++c) o[c] += acc[c];
}

}  // namespace

int fft_size_for(int G) {
  int np = 8;
  while (np < 2 * G) np <<= 1;
  return np;
}

static void fft3(float2* data, int64_t nChan, int NP, int inverse, const float2* tw, hipStream_t st) {
  const int64_t lines = nChan * (int64_t)NP * NP;
  const int bs = kFftBlock;
  const unsigned grid = (unsigned)((lines + bs - 1) / bs);
  const unsigned rowBlocks = (unsigned)(lines / NP);  // whole NP x NP blocks
  auto rows = [&](auto kern, int np) {
    hipLaunchKernelGGL(kern, dim3(rowBlocks), dim3(np), 0, st, data, inverse, tw);
  };
  if (NP == 32) rows(k_fft_rows<32>, 32);
  else if (NP == 64) rows(k_fft_rows<64>, 64);
  else if (NP == 16) rows(k_fft_rows<16>, 16);
  else rows(k_fft_rows<8>, 8);
  for (int axis = 1; axis < 3; ++axis) {
    if (NP == 32)
      hipLaunchKernelGGL(k_fft_lines<32>, dim3(grid), dim3(bs), 0, st, data, nChan, axis, inverse, tw);
    else if (NP == 64)
      hipLaunchKernelGGL(k_fft_lines<64>, dim3(grid), dim3(bs), 0, st, data, nChan, axis, inverse, tw);
    else if (NP == 16)
      hipLaunchKernelGGL(k_fft_lines<16>, dim3(grid), dim3(bs), 0, st, data, nChan, axis, inverse, tw);
    else
      hipLaunchKernelGGL(k_fft_lines<8>, dim3(grid), dim3(bs), 0, st, data, nChan, axis, inverse, tw);
  }
}

void launch_cells(const double* seg, int64_t ld, int n, const GridDev& g, int* cellOf,
                  unsigned* counts, unsigned* cursor, int* start, int* tmp, int* order,
                  hipStream_t st) {
  const int nc = g.G * g.G * g.G;
  (void)hipMemsetAsync(counts, 0, sizeof(unsigned) * nc, st);
  (void)hipMemsetAsync(cursor, 0, sizeof(unsigned) * nc, st);
  if (n > 0)
    hipLaunchKernelGGL(k_cell_assign, dim3((n + 255) / 256), dim3(256), 0, st, seg, ld, n, g, cellOf,
                       counts);
  hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(1024), 0, st, counts, nc, start);
  if (n > 0)
    hipLaunchKernelGGL(k_cell_scatter, dim3((n + 255) / 256), dim3(256), 0, st, cellOf, n, start,
                       cursor, tmp);
  hipLaunchKernelGGL(k_cell_rank, dim3(nc), dim3(64), 0, st, start, tmp, order);
}

void launch_moments(const double* seg, int64_t ld, const GridDev& g, const int* start,
                    const int* order, float* M, hipStream_t st) {
  const int nc = g.G * g.G * g.G;
  hipLaunchKernelGGL(k_cell_moments, dim3(nc), dim3(128), 0, st, seg, ld, g, start, order, M);
}

void launch_far_field(const float* M, const GridDev& g, const float2* What, const float2* tw,
                      float2* Mc, float2* Lc, float* L, hipStream_t st) {
  const int NP = fft_size_for(g.G);
  const int64_t vol = (int64_t)NP * NP * NP;
  const int64_t volH = (int64_t)NP * NP * (NP / 2 + 1);
  const int bs = 256;
  hipLaunchKernelGGL(k_pad_moments, dim3((unsigned)((kMomPairs * vol + bs - 1) / bs)), dim3(bs), 0,
                     st, M, g.G, NP, Mc);
  fft3(Mc, kMomPairs, NP, 0, tw, st);
  hipLaunchKernelGGL(k_spectral_apply, dim3((unsigned)volH), dim3(64), 0, st, What, Mc, NP, Lc);
  fft3(Lc, kLocPairs, NP, 1, tw, st);
  const int64_t nc = (int64_t)g.G * g.G * g.G;
  hipLaunchKernelGGL(k_unpad_local, dim3((unsigned)((kLocal * nc + bs - 1) / bs)), dim3(bs), 0, st,
                     Lc, g.G, NP, L);
}

void launch_operator_spectrum(const float* W, const GridDev& g, int rn, const float2* tw,
                              float2* scratch, float2* What, hipStream_t st) {
  const int NP = fft_size_for(g.G);
  const int64_t vol = (int64_t)NP * NP * NP;
  const int64_t volH = (int64_t)NP * NP * (NP / 2 + 1);
  const int bs = 256;
  hipLaunchKernelGGL(k_place_operator, dim3((unsigned)((kChanPairs * vol + bs - 1) / bs)), dim3(bs), 0,
                     st, W, g.G, NP, rn, scratch);
  fft3(scratch, kChanPairs, NP, 0, tw, st);
  hipLaunchKernelGGL(k_transpose_spectrum, dim3((unsigned)((kChan * volH + bs - 1) / bs)), dim3(bs), 0,
                     st, scratch, NP, What);
}

void launch_far_apply(const double* seg, int64_t ld, int n, const GridDev& g, const int* cellOf,
                      const float* L, double* out, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_far_apply, dim3((n + 255) / 256), dim3(256), 0, st, seg, ld, n, g, cellOf, L,
                     out);
}

void morton_ranks(int G, std::vector<int>& mrank, std::vector<int>& rankCell) {
  const int nc = G * G * G;
  auto spread = [](uint64_t x) {
    uint64_t r = 0;
    for (int b = 0; b < 10; ++b) r |= ((x >> b) & 1ull) << (3 * b);
    return r;
  };
  std::vector<std::pair<uint64_t, int>> code((size_t)nc);
  for (int c = 0; c < nc; ++c)
    code[(size_t)c] = {spread(c % G) | (spread((c / G) % G) << 1) | (spread(c / (G * G)) << 2), c};
  std::sort(code.begin(), code.end());
  mrank.assign((size_t)nc, 0);
  rankCell.assign((size_t)nc, 0);
  for (int r = 0; r < nc; ++r) {
    rankCell[(size_t)r] = code[(size_t)r].second;
    mrank[(size_t)code[(size_t)r].second] = r;
  }
}

void launch_near_list_count(const double* seg, int64_t ld, int n, const GridDev& g, int rn,
                            NearListDev& L, hipStream_t st) {
  const int nc = g.G * g.G * g.G;
  const int bs = 256;
  L.rn = rn;
  hipLaunchKernelGGL(k_morton_counts, dim3((nc + bs - 1) / bs), dim3(bs), 0, st, L.counts, L.rankCell,
                     nc, L.mcount);
  hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(1024), 0, st, L.mcount, nc, L.mstart);
  if (n > 0)
    hipLaunchKernelGGL(k_sort_gather, dim3((n + bs - 1) / bs), dim3(bs), 0, st, seg, ld, n, L.cellOf,
                       L.start, L.order, L.mrank, L.mstart, g.G, L.ldS, L.segS, L.key, L.perm);
  hipLaunchKernelGGL(k_sort_pad, dim3((unsigned)((L.ldS + bs - 1) / bs)), dim3(bs), 0, st, L.start, nc,
                     L.ldS, L.segS, L.key, L.perm);
  const int nbS = (int)(L.ldS / kTile);
  hipLaunchKernelGGL(k_block_boxes, dim3((nbS + bs - 1) / bs), dim3(bs), 0, st, L.key, L.start, nc, nbS,
                     L.box);
  hipLaunchKernelGGL(k_near_count_rows, dim3(nbS), dim3(64), 0, st, L.box, nbS, rn, L.rowCount);
  hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(1024), 0, st, L.rowCount, nbS, L.rowStart);
}

void launch_near_list_fill(const NearListDev& L, hipStream_t st) {
  const int nbS = (int)(L.ldS / kTile);
  hipLaunchKernelGGL(k_near_fill_rows, dim3(nbS), dim3(64), 0, st, L.box, nbS, L.rn, L.rowStart, L.list,
                     L.tileIdx);
  hipLaunchKernelGGL(k_near_count_cols, dim3(nbS), dim3(64), 0, st, L.tileIdx, nbS, L.colCount);
  hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(1024), 0, st, L.colCount, nbS, L.colStart);
  hipLaunchKernelGGL(k_near_fill_cols, dim3(nbS), dim3(64), 0, st, L.tileIdx, nbS, L.colStart,
                     L.colList);
}

void launch_near_reduce(const NearListDev& L, int nc, const double* rowPart, const double* colPart,
                        double* out, hipStream_t st) {
  const int bs = 256;
  hipLaunchKernelGGL(k_near_reduce, dim3((unsigned)((L.ldS + bs - 1) / bs)), dim3(bs), 0, st, rowPart,
                     colPart, L.rowStart, L.colStart, L.colList, L.perm, L.start, nc, out);
}

}  // namespace ddd
