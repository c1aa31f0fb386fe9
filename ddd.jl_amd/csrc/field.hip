// Per-segment, per-node and field-point kernels: self force, Peach-Koehler
// force from the FE correction stress, node-force gather, and the
// field-point stress sum.
#include "kernels.hpp"

namespace ddd {

namespace {

// _calcSelfForce (DislocationForces.jl:414-455).  Node 1 receives -f, node 2
// +f (:381-384).  rows[k] is the segment of output column k (or k itself
// when rows is null).
__global__ void k_self_force(const double* __restrict__ seg, int64_t ld,
                             const int64_t* __restrict__ rows, int n, SelfK K,
                             double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t i = rows ? rows[k] : k;
  if (i < 0 || seg[F_EXT * ld + i] != 0.0) return;
  const v3<double> n1 = V(seg[F_N1X * ld + i], seg[F_N1Y * ld + i], seg[F_N1Z * ld + i]);
  const v3<double> n2 = V(seg[F_N2X * ld + i], seg[F_N2Y * ld + i], seg[F_N2Z * ld + i]);
  const v3<double> b = V(seg[F_BX * ld + i], seg[F_BY * ld + i], seg[F_BZ * ld + i]);
  const v3<double> tv = n2 - n1;
  const double tvSq = dot(tv, tv);
  const double L = sqrt(tvSq);
  const double Linv = 1.0 / L;
  const v3<double> t = tv * Linv;
  const double La = sqrt(tvSq + K.aSq);
  const double bScrew = dot(t, b);
  const v3<double> bEdge = b - bScrew * t;
  const double bEdgeSq = dot(bEdge, bEdge);
  const double LaMa = La - K.a;
  const double tor = K.mu4pi * bScrew *
                     (K.nuOmNuInv * (ulog((La + L) / K.a) - 2.0 * LaMa * Linv) -
                      LaMa * LaMa / (2.0 * La * L));
  const double torCore = 2.0 * K.Ec * K.nuOmNuInv * bScrew;
  const double lonCore = (bScrew * bScrew + bEdgeSq * K.omNuInv) * K.Ec;
  const v3<double> f = (tor + torCore) * bEdge - lonCore * t;
  double* o = out + 6 * (int64_t)k;
  o[0] -= f.x; o[1] -= f.y; o[2] -= f.z;
  o[3] += f.x; o[4] += f.y; o[5] += f.z;
}

// calc_sigmaHat (DislocationForces.jl:92-192) at x: element by ceil/clamp,
// trilinear shape-function derivatives (FEMShapeFunction.jl:83-122), strain
// B U of the element's 24 uHat dofs, then C times strain; Voigt order
// (xx, yy, zz, xy, xz, yz).
__device__ void sigma_hat(const MeshDev& M, v3<double> x, double sv[6]) {
  const double wInv = 1.0 / M.w, hInv = 1.0 / M.h, dInv = 1.0 / M.d;
  auto cl = [](double v, int64_t hi) {
    double c = ceil(v);
    c = c < 1.0 ? 1.0 : c;
    c = c > (double)hi ? (double)hi : c;
    return (int64_t)c;
  };
  const int64_t ei = cl(x.x * wInv, M.mx), ej = cl(x.y * hInv, M.my), ek = cl(x.z * dInv, M.mz);
  const int64_t e = ei + (ej - 1) * M.mx + (ek - 1) * M.mx * M.my;
  const int64_t* ce = M.conn + 8 * (e - 1);
  const double* c1 = M.coord + 3 * (ce[0] - 1);
  const double* c7 = M.coord + 3 * (ce[6] - 1);
  const double ds1dx = 2.0 * wInv, ds2dy = 2.0 * hInv, ds3dz = 2.0 * dInv;
  const double s1 = (x.x - 0.5 * (c1[0] + c7[0])) * ds1dx;
  const double s2 = (x.y - 0.5 * (c1[1] + c7[1])) * ds2dy;
  const double s3 = (x.z - 0.5 * (c1[2] + c7[2])) * ds3dz;
  const double omx = 1 - s1, omy = 1 - s2, omz = 1 - s3, opx = 1 + s1, opy = 1 + s2, opz = 1 + s3;
  const double dN[3][8] = {
      {-omy * omz, omy * omz, opy * omz, -opy * omz, -omy * opz, omy * opz, opy * opz, -opy * opz},
      {-omx * omz, -opx * omz, opx * omz, omx * omz, -omx * opz, -opx * opz, opx * opz, omx * opz},
      {-omx * omy, -opx * omy, -opx * opy, -omx * opy, omx * omy, opx * omy, opx * opy, omx * opy}};
  double eps[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const double* u = M.uHat + 3 * (ce[q] - 1);
    const double gx = dN[0][q] / 8 * ds1dx, gy = dN[1][q] / 8 * ds2dy, gz = dN[2][q] / 8 * ds3dz;
    eps[0] += gx * u[0];
    eps[1] += gy * u[1];
    eps[2] += gz * u[2];
    eps[3] += gy * u[0] + gx * u[1];
    eps[4] += gz * u[0] + gx * u[2];
    eps[5] += gz * u[1] + gy * u[2];
  }
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    double a = 0;
#pragma unroll
    for (int c = 0; c < 6; ++c) a += M.C[r + 6 * c] * eps[c];
    sv[r] = a;
  }
}

// _calcPKForce (:292-300): pk = (sigmaHat(mid) b) x (n2 - n1), unnormalised
// tangent.  mode 0 writes pk to a 3 x n array (calcPKForce); mode 1 adds pk/2
// to both nodes of a 3 x 2 x n array (calcPKForce!, :266-269).
__global__ void k_pk_force(const double* __restrict__ seg, int64_t ld,
                           const int64_t* __restrict__ rows, int n, MeshDev M,
                           double* __restrict__ out, int mode) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t i = rows ? rows[k] : k;
  if (i < 0 || seg[F_EXT * ld + i] != 0.0) return;
  const v3<double> n1 = V(seg[F_N1X * ld + i], seg[F_N1Y * ld + i], seg[F_N1Z * ld + i]);
  const v3<double> n2 = V(seg[F_N2X * ld + i], seg[F_N2Y * ld + i], seg[F_N2Z * ld + i]);
  const v3<double> b = V(seg[F_BX * ld + i], seg[F_BY * ld + i], seg[F_BZ * ld + i]);
  const v3<double> t = n2 - n1;
  const v3<double> mid = V((n2.x + n1.x) / 2, (n2.y + n1.y) / 2, (n2.z + n1.z) / 2);
  double s[6];
  sigma_hat(M, mid, s);
  const v3<double> sb = V(s[0] * b.x + s[3] * b.y + s[4] * b.z, s[3] * b.x + s[1] * b.y + s[5] * b.z,
                          s[4] * b.x + s[5] * b.y + s[2] * b.z);
  const v3<double> pk = cross(sb, t);
  if (mode == 0) {
    double* o = out + 3 * (int64_t)k;
    o[0] += pk.x; o[1] += pk.y; o[2] += pk.z;
  } else {
    double* o = out + 6 * (int64_t)k;
    o[0] += pk.x / 2; o[1] += pk.y / 2; o[2] += pk.z / 2;
    o[3] += pk.x / 2; o[4] += pk.y / 2; o[5] += pk.z / 2;
  }
}

// calc_sigmaHat (:92-192) at P points x0 (3 x P): out holds 9 doubles per
// point, the reference's SMatrix{3,3} column by column (xx, xy, xz, xy, yy,
// yz, xz, yz, zz).
__global__ void k_sigma_hat(const double* __restrict__ x0, int64_t P, MeshDev M,
                            double* __restrict__ out) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= P) return;
  const v3<double> x = V(x0[3 * p], x0[3 * p + 1], x0[3 * p + 2]);
  double s[6];
  sigma_hat(M, x, s);
  const double m[9] = {s[0], s[3], s[4], s[3], s[1], s[5], s[4], s[5], s[2]};
  double* o = out + 9 * p;
#pragma unroll
  for (int q = 0; q < 9; ++q) o[q] = m[q];
}

// Field-point stress: SoA per segment of n1, n2, b (uploaded), t, the
// normalised tangent of :133, t × b, and three per-segment factors of the six
// stress components q = (xx, yy, zz, xy, xz, yz) = (a, b') (the component
// loop of :227-262): t_a t_b', the I13 coefficient -ν μ/(4π(1-ν)) (t⊗(t×b))_q
// and the first term a² μ/(8π) (t⊗(t×b))_q of I15.  They depend on the
// segment only, so the pair kernel reads them instead of recomputing them at
// every point; every value is the same IEEE expression either way.
enum { S_N1 = 0, S_N2 = 3, S_B = 6, S_T = 9, S_TXB = 12, S_TMT = 15, S_I13 = 21, S_I15 = 27,
       S_COUNT = 33 };
static_assert(S_COUNT == kSigmaFields, "sigma segment SoA");

constexpr int kSigQa[6] = {0, 1, 2, 0, 0, 1}, kSigQb[6] = {0, 1, 2, 1, 2, 2};

__global__ void k_prep_sigma_segments(double* __restrict__ s, int64_t ld, int n, SigmaK K) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dx = s[(S_N2 + 0) * ld + i] - s[(S_N1 + 0) * ld + i];
  const double dy = s[(S_N2 + 1) * ld + i] - s[(S_N1 + 1) * ld + i];
  const double dz = s[(S_N2 + 2) * ld + i] - s[(S_N1 + 2) * ld + i];
  const double nn = __dsqrt_rn(__dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)),
                                         __dmul_rn(dz, dz)));
  const v3<double> t = V(__ddiv_rn(dx, nn), __ddiv_rn(dy, nn), __ddiv_rn(dz, nn));
  s[(S_T + 0) * ld + i] = t.x;
  s[(S_T + 1) * ld + i] = t.y;
  s[(S_T + 2) * ld + i] = t.z;
  const v3<double> b = V(s[(S_B + 0) * ld + i], s[(S_B + 1) * ld + i], s[(S_B + 2) * ld + i]);
  const v3<double> txb = cross(t, b);
  s[(S_TXB + 0) * ld + i] = txb.x;
  s[(S_TXB + 1) * ld + i] = txb.y;
  s[(S_TXB + 2) * ld + i] = txb.z;
  const double tt[3] = {t.x, t.y, t.z}, xb[3] = {txb.x, txb.y, txb.z};
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const int a = kSigQa[q], bq = kSigQb[q];
    const double tmtxb = q < 3 ? 2.0 * tt[a] * xb[a] : tt[a] * xb[bq] + tt[bq] * xb[a];
    s[(S_TMT + q) * ld + i] = tt[a] * tt[bq];
    s[(S_I13 + q) * ld + i] = -K.nuMu4pinu * tmtxb;
    s[(S_I15 + q) * ld + i] = K.a2mu8pi * tmtxb;
  }
}

__global__ void k_gather_sigma_segments(const int64_t* __restrict__ segIdx, int64_t L,
                                        int64_t numSeg, const double* __restrict__ coord,
                                        int64_t coordCols, const double* __restrict__ bVec,
                                        int64_t bVecCols, const int64_t* __restrict__ idx, int m,
                                        double* __restrict__ s, int64_t ld) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m) return;
  const int64_t i = idx ? idx[q] - 1 : q;
  bool ok = i >= 0 && i < numSeg;
  const int64_t lk = ok ? segIdx[i] : 0, a = ok ? segIdx[i + L] : 0, b = ok ? segIdx[i + 2 * L] : 0;
  ok = ok && lk >= 1 && lk <= bVecCols && a >= 1 && a <= coordCols && b >= 1 && b <= coordCols;
  const double nan = __longlong_as_double(0x7ff8000000000000ll);
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    s[(S_N1 + e) * ld + q] = ok ? coord[3 * (a - 1) + e] : nan;
    s[(S_N2 + e) * ld + q] = ok ? coord[3 * (b - 1) + e] : nan;
    s[(S_B + e) * ld + q] = ok ? bVec[3 * (lk - 1) + e] : nan;
  }
}

constexpr int kSigBlock = 256;
constexpr int kSigTile = 64;

// Stress of segments [c * chunk, (c + 1) * chunk) at one point per thread,
// summed in segment order (the reference's order for each point).  Each
// staged 64-segment tile is evaluated with the ranged division and square
// root of fp64_ranged.hpp when the workgroup's points, the tile's end points
// and Burgers vectors and a^2 are in their safe range (the same bits), else
// with the general IEEE sequences; a workgroup with a non-finite sum
// evaluates its whole chunk again with the general ones.
__global__ __launch_bounds__(kSigBlock) void k_sigma_tilde(const double* __restrict__ sseg,
                                                           int64_t sld, int nseg, int chunk,
                                                           const double* __restrict__ x0,
                                                           int64_t P, SigmaK K,
                                                           double* __restrict__ partial) {
  __shared__ double sh[S_COUNT][kSigTile];
  const int64_t p = blockIdx.x * (int64_t)kSigBlock + threadIdx.x;
  const int c = blockIdx.y;
  const int s0 = c * chunk;
  const int s1 = min(nseg, s0 + chunk);
  const bool live = p < P;
  const v3<double> x = live ? V(x0[3 * p], x0[3 * p + 1], x0[3 * p + 2]) : V(0.0, 0.0, 0.0);
  const bool pointOk = ranged_value(x.x) && ranged_value(x.y) && ranged_value(x.z);
  const bool aOk = K.a2 >= 0x1p-120 && K.a2 <= 0x1p120;
  double sg[6] = {0, 0, 0, 0, 0, 0};

  // the stress of staged segments [0, cnt) at x, added to sg
  auto tile = [&]<bool RANGED>(int cnt) {
    for (int l = 0; l < cnt; ++l) {
      const v3<double> n1 = V(sh[S_N1][l], sh[S_N1 + 1][l], sh[S_N1 + 2][l]);
      const v3<double> n2 = V(sh[S_N2][l], sh[S_N2 + 1][l], sh[S_N2 + 2][l]);
      const v3<double> b = V(sh[S_B][l], sh[S_B + 1][l], sh[S_B + 2][l]);
      const v3<double> t = V(sh[S_T][l], sh[S_T + 1][l], sh[S_T + 2][l]);
      // DislocationFEM.jl:142-299 (per-segment factors from k_prep_sigma_segments)
      const v3<double> R = x - n1;
      const double Rdt = dot(R, t);
      const v3<double> nd = R - Rdt * t;
      const double d2 = dot(nd, nd);
      const double sa = -Rdt;
      const double sb = -dot(x - n2, t);
      const double a2_d2 = K.a2 + d2;
      double a2d2inv, Ra, Rainv;
      if constexpr (RANGED) {
        a2d2inv = div_ranged(1.0, a2_d2);
        Ra = sqrt_ranged(a2_d2 + sa * sa);
        Rainv = div_ranged(1.0, Ra);
      } else {
        a2d2inv = 1.0 / a2_d2;
        Ra = sqrt(a2_d2 + sa * sa);
        Rainv = 1.0 / Ra;
      }
      double Ra3inv = Rainv * Rainv * Rainv;
      double sRa3inv = sa * Ra3inv;
      const double s03a = sa * Rainv * a2d2inv;
      const double s13a = -Rainv;
      const double s05a = (2.0 * s03a + sRa3inv) * a2d2inv;
      const double s15a = -Ra3inv;
      const double s25a = s03a - sRa3inv;
      if constexpr (RANGED) {
        Ra = sqrt_ranged(a2_d2 + sb * sb);
        Rainv = div_ranged(1.0, Ra);
      } else {
        Ra = sqrt(a2_d2 + sb * sb);
        Rainv = 1.0 / Ra;
      }
      Ra3inv = Rainv * Rainv * Rainv;
      sRa3inv = sb * Ra3inv;
      const double s03b = sb * Rainv * a2d2inv;
      const double s13b = -Rainv;
      const double s05b = (2.0 * s03b + sRa3inv) * a2d2inv;
      const double s15b = -Ra3inv;
      const double s25b = s03b - sRa3inv;
      const double s03 = s03b - s03a, s13 = s13b - s13a, s05 = s05b - s05a, s15 = s15b - s15a,
                   s25 = s25b - s25a;
      const v3<double> dxb = cross(nd, b);
      const double common = K.mu4pinu * dot(dxb, t);
      const double tt[3] = {t.x, t.y, t.z}, dd[3] = {nd.x, nd.y, nd.z};
      const double xb[3] = {sh[S_TXB][l], sh[S_TXB + 1][l], sh[S_TXB + 2][l]};
      const double db[3] = {dxb.x, dxb.y, dxb.z};
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        const int a = kSigQa[q], bq = kSigQb[q];
        const bool dg = q < 3;
        const double dmd = dd[a] * dd[bq];
        const double tmd = dg ? 2.0 * tt[a] * dd[a] : tt[a] * dd[bq] + tt[bq] * dd[a];
        const double dmtxb = dg ? 2.0 * dd[a] * xb[a] : dd[a] * xb[bq] + dd[bq] * xb[a];
        const double tmdxb = dg ? 2.0 * tt[a] * db[a] : tt[a] * db[bq] + tt[bq] * db[a];
        const double I03 = (dg ? common : 0.0) + K.mu4pinu * dmtxb - K.mu4pi * tmdxb;
        const double I13 = sh[S_I13 + q][l];
        const double I05 = common * ((dg ? K.a2 : 0.0) + dmd) - K.a2mu8pi * tmdxb;
        const double I15 = sh[S_I15 + q][l] - common * tmd;
        const double I25 = common * sh[S_TMT + q][l];
        sg[q] += I03 * s03 + I13 * s13 + I05 * s05 + I15 * s15 + I25 * s25;
      }
    }
  };

  // every tile of the chunk; ranged where allowed (and not forced off)
  auto sweep = [&](bool allowRanged) {
    for (int base = s0; base < s1; base += kSigTile) {
      const int cnt = min(kSigTile, s1 - base);
      __syncthreads();
      bool bad = !pointOk;
      for (int q = threadIdx.x; q < S_COUNT * kSigTile; q += kSigBlock) {
        const int f = q / kSigTile, l = q % kSigTile;
        const double v = (l < cnt) ? sseg[f * sld + base + l] : 0.0;
        sh[f][l] = v;
        if (f < S_T) bad = bad || !ranged_value(v);  // n1, n2, b
      }
      const bool tileBad = __syncthreads_or(bad);
      if (allowRanged && aOk && !tileBad) tile.template operator()<true>(cnt);
      else tile.template operator()<false>(cnt);
    }
  };

  sweep(true);
  double tot = 0.0;
#pragma unroll
  for (int q = 0; q < 6; ++q) tot = tot + sg[q];
  if (__syncthreads_or(!(tot - tot == 0.0))) {  // inf or NaN in a workgroup's sum
#pragma unroll
    for (int q = 0; q < 6; ++q) sg[q] = 0.0;
    sweep(false);
  }
  if (live) {
    double* o = partial + ((int64_t)c * P + p) * 6;
#pragma unroll
    for (int q = 0; q < 6; ++q) o[q] = sg[q];
  }
}

__global__ void k_sigma_reduce(const double* __restrict__ partial, int nChunk, int64_t P,
                               double* __restrict__ sigma, int accumulate) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= 6 * P) return;
  double a = 0.0;
  for (int c = 0; c < nChunk; ++c) a += partial[(int64_t)c * 6 * P + e];
  sigma[e] = accumulate ? sigma[e] + a : a;
}

// dlnMobility! node-force gather (DislocationMobility.jl:220-255).
__global__ void k_node_gather(const int64_t* __restrict__ conn, int64_t ldConn, int64_t numNode,
                              const int64_t* __restrict__ links, const double* __restrict__ coord,
                              const double* __restrict__ segForce, double sigmaPN,
                              double* __restrict__ nodeForce) {
  const int64_t nd = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (nd >= numNode) return;
  const int64_t* cn = conn + ldConn * nd;
  double f0 = 0, f1 = 0, f2 = 0;
  const int64_t cnt = cn[0];
  for (int64_t c = 1; c <= cnt; ++c) {
    const int64_t link = cn[2 * c - 1], col = cn[2 * c];
    const int64_t node2 = links[2 * (link - 1) + (2 - col)];
    const double tx = coord[3 * (node2 - 1)] - coord[3 * nd];
    const double ty = coord[3 * (node2 - 1) + 1] - coord[3 * nd + 1];
    const double tz = coord[3 * (node2 - 1) + 2] - coord[3 * nd + 2];
    const double nT = sqrt(tx * tx + ty * ty + tz * tz);
    if (nT < 2.220446049250313e-16) continue;
    const double* sf = segForce + 6 * (link - 1) + 3 * (col - 1);
    double gx = sf[0], gy = sf[1], gz = sf[2];
    if (sqrt(gx * gx + gy * gy + gz * gz) * (1.0 / nT) < sigmaPN) gx = gy = gz = 0.0;
    f0 += gx;
    f1 += gy;
    f2 += gz;
  }
  nodeForce[3 * nd] = f0;
  nodeForce[3 * nd + 1] = f1;
  nodeForce[3 * nd + 2] = f2;
}

// BCC mobility (dlnMobility / dlnMobility! for mobBCC,
// DislocationMobility.jl:16-171,186-341), one thread per node: the node force
// (the gather above, :56-80), the drag tensor of the node's links (:82-146)
// and the velocity solving drag v = f (:148) by Cramer's rule (the method
// of the reference's 3x3 StaticArrays solve).  Node k of the list is
// idx[k] (1-based) or k + 1 without a list; results go to column k.
struct DragDev {
  double edge, screw, climb, line;
};

__global__ void __launch_bounds__(256)
k_mobility_bcc(const int64_t* __restrict__ conn, int64_t ldConn, const int64_t* __restrict__ links,
               const double* __restrict__ coord, const double* __restrict__ bVec,
               const double* __restrict__ segForce, const int64_t* __restrict__ idx, int64_t n,
               DragDev drag, double sigmaPN, const int32_t* __restrict__ label,
               double* __restrict__ nodeForce, double* __restrict__ nodeVel) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double eps = 2.220446049250313e-16;
  const double sqrtEps = 1.4901161193847656e-08;  // sqrt(eps(Float64)), exact
  const int64_t nd = idx ? idx[k] - 1 : k;
  const int64_t* cn = conn + ldConn * nd;
  const double x0 = coord[3 * nd], y0 = coord[3 * nd + 1], z0 = coord[3 * nd + 2];
  // The full 3x3 tensor, entry [i][j] accumulated as the reference forms
  // it ((s t_i) t_j): its two halves differ by rounding, and with drag
  // ratios up to ~1e10 that rounding reaches the velocity.
  double D[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double f0 = 0, f1 = 0, f2 = 0;
  const int64_t cnt = cn[0];
  for (int64_t c = 1; c <= cnt; ++c) {
    const int64_t link = cn[2 * c - 1], col = cn[2 * c];
    const int64_t node2 = links[2 * (link - 1) + (2 - col)];
    double tx = coord[3 * (node2 - 1)] - x0;
    double ty = coord[3 * (node2 - 1) + 1] - y0;
    double tz = coord[3 * (node2 - 1) + 2] - z0;
    const double nT = sqrt(tx * tx + ty * ty + tz * tz);
    if (nT < eps) continue;
    const double tN = 1.0 / nT;
    tx = tx * tN;
    ty = ty * tN;
    tz = tz * tN;
    const double* sf = segForce + 6 * (link - 1) + 3 * (col - 1);
    double gx = sf[0], gy = sf[1], gz = sf[2];
    if (sqrt(gx * gx + gy * gy + gz * gz) * tN < sigmaPN) gx = gy = gz = 0.0;
    f0 += gx;
    f1 += gy;
    f2 += gz;

    double bx = bVec[3 * (link - 1)], by = bVec[3 * (link - 1) + 1], bz = bVec[3 * (link - 1) + 2];
    const double nB = sqrt(bx * bx + by * by + bz * bz);
    // <111>-family test on |b_i| / min |b_j|; a zero component gives
    // Inf/NaN ratios that match no family.
    const double ax = fabs(bx), ay = fabs(by), az = fabs(bz);
    const double myz = az < ay ? az : ay;
    const double mn = myz < ax ? myz : ax;
    int c1 = 0, c2 = 0, c3 = 0;
    {
      const double r[3] = {ax / mn, ay / mn, az / mn};
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        c1 += fabs(1.0 - r[q]) < sqrtEps;
        c2 += fabs(2.0 - r[q]) < sqrtEps;
        c3 += fabs(3.0 - r[q]) < sqrtEps;
      }
    }
    const bool screwFamily =
        (1.0 - nB < sqrtEps) && (c1 == 3 || (c1 == 2 && c2 == 1) || (c1 == 1 && c2 == 1 && c3 == 1));
    bx = bx / nB;
    by = by / nB;
    bz = bz / nB;
    const double w = nB * nT / 2.0;
    const double base = screwFamily ? drag.screw : drag.climb;
    const double tv[3] = {tx, ty, tz};
    // w (base I + (line - base) t t)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        D[i][j] += w * ((i == j ? base : 0.0) + (drag.line - base) * tv[i] * tv[j]);
    if (!screwFamily) continue;
    const double tDb = tx * bx + ty * by + tz * bz;
    const double cosSq = tDb * tDb;
    const double sinSq = 1.0 - cosSq;
    if (sinSq < eps) continue;
    const double sq = sqrt(sinSq);
    const double px = (by * tz - bz * ty) / sq;
    const double py = (bz * tx - bx * tz) / sq;
    const double pz = (bx * ty - by * tx) / sq;
    const double qx = py * tz - pz * ty;
    const double qy = pz * tx - px * tz;
    const double qz = px * ty - py * tx;
    const double glide =
        1.0 / sqrt(sinSq / (drag.edge * drag.edge) + cosSq / (drag.screw * drag.screw));
    const double climbS = sqrt(sinSq * (drag.climb * drag.climb) + cosSq * (drag.screw * drag.screw));
    const double gq = glide - drag.screw, cp = climbS - drag.screw;
    const double pv[3] = {px, py, pz}, qv[3] = {qx, qy, qz};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) D[i][j] += w * (gq * qv[i] * qv[j] + cp * pv[i] * pv[j]);
  }
  // Columns c0, c1, c2 of the tensor; the inverse's rows are c1 x c2,
  // c2 x c0, c0 x c1 over det = c0 . (c1 x c2).
  const double c0x = D[0][0], c0y = D[1][0], c0z = D[2][0];
  const double c1x = D[0][1], c1y = D[1][1], c1z = D[2][1];
  const double c2x = D[0][2], c2y = D[1][2], c2z = D[2][2];
  const double r0x = c1y * c2z - c1z * c2y, r0y = c1z * c2x - c1x * c2z, r0z = c1x * c2y - c1y * c2x;
  const double r1x = c2y * c0z - c2z * c0y, r1y = c2z * c0x - c2x * c0z, r1z = c2x * c0y - c2y * c0x;
  const double r2x = c0y * c1z - c0z * c1y, r2y = c0z * c1x - c0x * c1z, r2z = c0x * c1y - c0y * c1x;
  const double det = c0x * r0x + c0y * r0y + c0z * r0z;
  nodeForce[3 * k] = f0;
  nodeForce[3 * k + 1] = f1;
  nodeForce[3 * k + 2] = f2;
  double v0 = (r0x * f0 + r0y * f1 + r0z * f2) / det;
  double v1 = (r1x * f0 + r1y * f1 + r1z * f2) / det;
  double v2 = (r2x * f0 + r2y * f1 + r2z * f2) / det;
  if (label) {
    // deriv! (IntegrationProcessing.jl:29-32): internal fixed, surface fixed
    // and external nodes (nodeTypeDln 2, 4, 5) do not move.
    const int32_t lb = label[nd];
    if (lb == 2 || lb == 4 || lb == 5) v0 = v1 = v2 = 0.0;
  }
  nodeVel[3 * k] = v0;
  nodeVel[3 * k + 1] = v1;
  nodeVel[3 * k + 2] = v2;
}

}  // namespace

void launch_self_force(const double* seg, int64_t ld, const int64_t* rows, int n, const SelfK& K,
                       double* out, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_self_force, dim3((n + 255) / 256), dim3(256), 0, st, seg, ld, rows, n, K,
                     out);
}

void launch_pk_force(const double* seg, int64_t ld, const int64_t* rows, int n, const MeshDev& M,
                     double* out, int mode, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_pk_force, dim3((n + 255) / 256), dim3(256), 0, st, seg, ld, rows, n, M,
                     out, mode);
}

void launch_sigma_hat(const double* x0, int64_t P, const MeshDev& M, double* out, hipStream_t st) {
  if (P <= 0) return;
  hipLaunchKernelGGL(k_sigma_hat, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, x0, P, M, out);
}

void launch_prep_sigma_segments(double* sseg, int64_t ld, int n, const SigmaK& K, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_prep_sigma_segments, dim3((n + 255) / 256), dim3(256), 0, st, sseg, ld, n, K);
}

void launch_gather_sigma_segments(const int64_t* segIdx, int64_t segIdxLd, int64_t numSeg,
                                  const double* coord, int64_t coordCols, const double* bVec,
                                  int64_t bVecCols, const int64_t* idx, int m, double* sseg,
                                  int64_t ld, hipStream_t st) {
  if (m <= 0) return;
  hipLaunchKernelGGL(k_gather_sigma_segments, dim3((m + 255) / 256), dim3(256), 0, st, segIdx,
                     segIdxLd, numSeg, coord, coordCols, bVec, bVecCols, idx, m, sseg, ld);
}

int sigma_chunks_for(int64_t P, int nseg) {
  // Enough blocks to fill the device several times over, each chunk still
  // long enough to amortise its LDS staging.
  const int64_t pblocks = (P + kSigBlock - 1) / kSigBlock;
  int c = 1;
  while (pblocks * c < 4096 && nseg / (2 * c) >= 2 * kSigTile) c *= 2;
  return c;
}

void launch_sigma_tilde(const double* sseg, int64_t sld, int nseg, const double* x0, int64_t P,
                        const SigmaK& K, double* partial, int nChunk, hipStream_t st) {
  if (P <= 0) return;
  const int chunk = (nseg + nChunk - 1) / nChunk;
  dim3 grid((unsigned)((P + kSigBlock - 1) / kSigBlock), (unsigned)nChunk);
  hipLaunchKernelGGL(k_sigma_tilde, grid, dim3(kSigBlock), 0, st, sseg, sld, nseg, chunk, x0, P, K,
                     partial);
}

void launch_sigma_reduce(const double* partial, int nChunk, int64_t P, double* sigma,
                         int accumulate, hipStream_t st) {
  if (P <= 0) return;
  const int64_t n = 6 * P;
  hipLaunchKernelGGL(k_sigma_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, partial,
                     nChunk, P, sigma, accumulate);
}

void launch_node_gather(const int64_t* conn, int64_t ldConn, int64_t numNode,
                        const int64_t* links, const double* coord, const double* segForce,
                        double sigmaPN, double* nodeForce, hipStream_t st) {
  if (numNode <= 0) return;
  hipLaunchKernelGGL(k_node_gather, dim3((unsigned)((numNode + 255) / 256)), dim3(256), 0, st, conn,
                     ldConn, numNode, links, coord, segForce, sigmaPN, nodeForce);
}

void launch_mobility_bcc(const int64_t* conn, int64_t ldConn, const int64_t* links,
                         const double* coord, const double* bVec, const double* segForce,
                         const int64_t* idx, int64_t n, const double drag[4], double sigmaPN,
                         const int32_t* label, double* nodeForce, double* nodeVel,
                         hipStream_t st) {
  if (n <= 0) return;
  const DragDev d{drag[0], drag[1], drag[2], drag[3]};
  hipLaunchKernelGGL(k_mobility_bcc, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, conn,
                     ldConn, links, coord, bVec, segForce, idx, n, d, sigmaPN, label, nodeForce,
                     nodeVel);
}

}  // namespace ddd
