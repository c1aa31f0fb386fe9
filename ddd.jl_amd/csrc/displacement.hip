// Dislocation displacements at field points: the Barnett-triangle sum of
// calc_uTilde (src/Processing/DislocationFEM.jl:355-707; Bromage & Tarleton,
// MSMSE 26 (2018) 085007).  Built without FMA contraction: every operation
// rounds as the reference's expressions, written in the same order.
//
// Each triangle (A, B, C, b) is either "full" (an internal segment AB closed
// at the volume's centroid C: solid angle plus the material line terms,
// :602-670) or "solid" (one of the closure triangles of an external segment:
// the solid-angle term alone, :671-707).  One lane per point; triangles are
// staged through LDS in tiles with their per-triangle quantities (unit edges,
// b x edge, CA x AB) computed once per tile; a grid dimension splits the
// triangles into chunks whose partial sums a second kernel adds in fixed
// order.
#include "kernels.hpp"

namespace ddd {

namespace {

constexpr int kUBlock = 256;
constexpr int kUTile = 64;

enum UShared {
  U_A = 0, U_B = 3, U_C = 6, U_BV = 9, U_AB = 12, U_BC = 15, U_CA = 18, U_BAB = 21, U_BBC = 24,
  U_BCA = 27, U_CAXAB = 30, U_KIND = 33, U_COUNT = 34
};

struct dv3 {
  double x, y, z;
};
__device__ inline dv3 mk3(double a, double b, double c) { return dv3{a, b, c}; }
__device__ inline dv3 sub(const dv3& a, const dv3& b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline dv3 add(const dv3& a, const dv3& b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline dv3 scl(double s, const dv3& a) { return mk3(s * a.x, s * a.y, s * a.z); }
__device__ inline double dot3(const dv3& a, const dv3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline dv3 cross3(const dv3& a, const dv3& b) {
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// safeNorm (Misc.jl:289-299): (|x|, x / |x|), or (eps, 0) at or below eps.
__device__ inline dv3 safe_norm(const dv3& x, double& n) {
  const double eps = 2.220446049250313e-16;
  n = sqrt(x.x * x.x + x.y * x.y + x.z * x.z);
  if (n > eps) return mk3(x.x / n, x.y / n, x.z / n);
  n = eps;
  return mk3(0.0, 0.0, 0.0);
}
__device__ inline double clamp1(double x) { return x < -1.0 ? -1.0 : (x > 1.0 ? 1.0 : x); }
__device__ inline double sgn(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }

__device__ inline dv3 ld3(const double (*sh)[kUTile], int f, int l) {
  return mk3(sh[f][l], sh[f + 1][l], sh[f + 2][l]);
}
__device__ inline void st3(double (*sh)[kUTile], int f, int l, const dv3& v) {
  sh[f][l] = v.x;
  sh[f + 1][l] = v.y;
  sh[f + 2][l] = v.z;
}

__global__ __launch_bounds__(kUBlock) void k_u_tilde(const double* __restrict__ tri, int64_t tld,
                                                     int ntri, int chunk,
                                                     const double* __restrict__ pts, int64_t P,
                                                     double omNuInv8pi, double om2NuOmNuInv8pi,
                                                     double* __restrict__ partial) {
  __shared__ double sh[U_COUNT][kUTile];
  const int64_t p = blockIdx.x * (int64_t)kUBlock + threadIdx.x;
  const int c = blockIdx.y;
  const int t0 = c * chunk;
  const int t1 = min(ntri, t0 + chunk);
  const bool live = p < P;
  const dv3 x = live ? mk3(pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]) : mk3(0.0, 0.0, 0.0);
  const double fourPi = 4.0 * 3.141592653589793;
  double u0 = 0.0, u1 = 0.0, u2 = 0.0;
  for (int base = t0; base < t1; base += kUTile) {
    const int cnt = min(kUTile, t1 - base);
    __syncthreads();
    if (threadIdx.x < cnt) {
      const int l = threadIdx.x;
      const int64_t q = base + l;
      const dv3 A = mk3(tri[0 * tld + q], tri[1 * tld + q], tri[2 * tld + q]);
      const dv3 B = mk3(tri[3 * tld + q], tri[4 * tld + q], tri[5 * tld + q]);
      const dv3 C = mk3(tri[6 * tld + q], tri[7 * tld + q], tri[8 * tld + q]);
      const dv3 b = mk3(tri[9 * tld + q], tri[10 * tld + q], tri[11 * tld + q]);
      double dummy;
      const dv3 AB = safe_norm(sub(B, A), dummy);
      const dv3 BC = safe_norm(sub(C, B), dummy);
      const dv3 CA = safe_norm(sub(A, C), dummy);
      st3(sh, U_A, l, A);
      st3(sh, U_B, l, B);
      st3(sh, U_C, l, C);
      st3(sh, U_BV, l, b);
      st3(sh, U_AB, l, AB);
      st3(sh, U_BC, l, BC);
      st3(sh, U_CA, l, CA);
      st3(sh, U_BAB, l, cross3(b, AB));
      st3(sh, U_BBC, l, cross3(b, BC));
      st3(sh, U_BCA, l, cross3(b, CA));
      st3(sh, U_CAXAB, l, cross3(CA, AB));
      sh[U_KIND][l] = tri[12 * tld + q];
    }
    __syncthreads();
    if (!live) continue;
    for (int l = 0; l < cnt; ++l) {
      double ra, rb, rc;
      const dv3 RA = safe_norm(sub(ld3(sh, U_A, l), x), ra);
      const dv3 RB = safe_norm(sub(ld3(sh, U_B, l), x), rb);
      const dv3 RC = safe_norm(sub(ld3(sh, U_C, l), x), rc);
      // solid angle by l'Huilier's theorem (DislocationFEM.jl:648-657)
      const double ab = clamp1(dot3(RA, RB)), bc = clamp1(dot3(RB, RC)), ca = clamp1(dot3(RC, RA));
      const double ta = acos(ab), tb = acos(bc), tc = acos(ca);
      const double th = (ta + tb + tc) / 2.0;
      double f = tan(th / 2.0) * tan((th - ta) / 2.0) * tan((th - tb) / 2.0) * tan((th - tc) / 2.0);
      f = (f < 0.0) ? 0.0 : f;  // max(f, 0); NaN stays NaN
      const double w = -sgn(dot3(RA, ld3(sh, U_CAXAB, l))) * (4.0 * atan(sqrt(f)));
      const dv3 b = ld3(sh, U_BV, l);
      const double s0 = (-b.x * w) / fourPi, s1 = (-b.y * w) / fourPi, s2 = (-b.z * w) / fourPi;
      if (sh[U_KIND][l] == 0.0) {
        u0 += s0;
        u1 += s1;
        u2 += s2;
        continue;
      }
      // line terms f and g (:630-646)
      const dv3 AB = ld3(sh, U_AB, l), BC = ld3(sh, U_BC, l), CA = ld3(sh, U_CA, l);
      const double lAB = log((rb / ra) * (1.0 + dot3(RB, AB)) / (1.0 + dot3(RA, AB)));
      const double lBC = log((rc / rb) * (1.0 + dot3(RC, BC)) / (1.0 + dot3(RB, BC)));
      const double lCA = log((ra / rc) * (1.0 + dot3(RA, CA)) / (1.0 + dot3(RC, CA)));
      const dv3 fAB = scl(lAB, ld3(sh, U_BAB, l));
      const dv3 fBC = scl(lBC, ld3(sh, U_BBC, l));
      const dv3 fCA = scl(lCA, ld3(sh, U_BCA, l));
      const double gA = dot3(b, cross3(RA, RB)), gB = dot3(b, cross3(RB, RC)),
                   gC = dot3(b, cross3(RC, RA));
      const dv3 sAB = add(RA, RB), sBC = add(RB, RC), sCA = add(RC, RA);
      const dv3 gAB = mk3((gA * sAB.x) / (1.0 + ab), (gA * sAB.y) / (1.0 + ab), (gA * sAB.z) / (1.0 + ab));
      const dv3 gBC = mk3((gB * sBC.x) / (1.0 + bc), (gB * sBC.y) / (1.0 + bc), (gB * sBC.z) / (1.0 + bc));
      const dv3 gCA = mk3((gC * sCA.x) / (1.0 + ca), (gC * sCA.y) / (1.0 + ca), (gC * sCA.z) / (1.0 + ca));
      const dv3 fs = add(add(fAB, fBC), fCA);
      const dv3 gs = add(add(gAB, gBC), gCA);
      u0 += (s0 - om2NuOmNuInv8pi * fs.x) + omNuInv8pi * gs.x;
      u1 += (s1 - om2NuOmNuInv8pi * fs.y) + omNuInv8pi * gs.y;
      u2 += (s2 - om2NuOmNuInv8pi * fs.z) + omNuInv8pi * gs.z;
    }
  }
  if (live) {
    double* out = partial + (int64_t)c * 3 * P;
    out[3 * p] = u0;
    out[3 * p + 1] = u1;
    out[3 * p + 2] = u2;
  }
}

__global__ void k_chunk_sum(const double* __restrict__ partial, int nChunk, int64_t count,
                            double* __restrict__ out, int accumulate) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= count) return;
  double a = 0.0;
  for (int c = 0; c < nChunk; ++c) a += partial[(int64_t)c * count + e];
  out[e] = accumulate ? out[e] + a : a;
}

}  // namespace

int u_tilde_chunks_for(int64_t P, int ntri) {
  const int64_t pblocks = (P + kUBlock - 1) / kUBlock;
  int c = 1;
  while (pblocks * c < 4096 && ntri / (2 * c) >= 2 * kUTile) c *= 2;
  return c;
}

void launch_u_tilde(const double* tri, int64_t tld, int ntri, const double* pts, int64_t P,
                    double omNuInv8pi, double om2NuOmNuInv8pi, double* partial, int nChunk,
                    hipStream_t st) {
  if (P <= 0) return;
  const int chunk = (ntri + nChunk - 1) / nChunk;
  dim3 grid((unsigned)((P + kUBlock - 1) / kUBlock), (unsigned)nChunk);
  hipLaunchKernelGGL(k_u_tilde, grid, dim3(kUBlock), 0, st, tri, tld, ntri, chunk, pts, P,
                     omNuInv8pi, om2NuOmNuInv8pi, partial);
}

void launch_chunk_sum(const double* partial, int nChunk, int64_t count, double* out,
                      int accumulate, hipStream_t st) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_chunk_sum, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st,
                     partial, nChunk, count, out, accumulate);
}

}  // namespace ddd
