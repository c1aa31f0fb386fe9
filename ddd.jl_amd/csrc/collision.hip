// Collision detection on the device: detectCollision
// (src/Processing/DislocationTopologyCollision.jl:1-222) with minimumDistance
// (src/Misc/Misc.jl:314-367).  Built without FMA contraction so that every
// distance, rate and 1/distance rounds as the reference's expressions.
//
// The reference scans hinges (two links at one mobile node) and then, if no
// hinge collides, all unordered pairs of links that share no node, keeping the
// candidate with the largest 1/distSq (the first one in loop order on ties)
// and returning L1, L2 of the last pair it evaluated.  Here both scans are
// reductions with that order made explicit: a candidate is (key, loop
// position), compared by key, then by position; the last evaluated pair is
// the largest position; its L1, L2 come with it (hinges) or are computed by
// the final reduction (pairs).
//
// The pair scan visits links in Morton order of their midpoints (radix
// sort), 64 to a block, and skips every tile of two blocks whose bounding
// boxes are farther apart than the collision distance: no pair in it can be
// a candidate.  Each pair keeps the reference's orientation (lower link
// index first) for its arithmetic and its loop position.  The last
// evaluated pair does not depend on distances; it is found per row by
// scanning its columns from the end (k_collision_last).
#include <hipcub/hipcub.hpp>

#include "kernels.hpp"

namespace ddd {

namespace {

constexpr double kEps = 2.220446049250313e-16;

struct MinDist {
  double distSq, dDistSqDt, L1, L2;
};

__device__ inline double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// minimumDistance (Misc.jl:314-367), the reference's branches and clamps.
__device__ MinDist min_distance(const double* x0, const double* x1, const double* y0,
                                const double* y1, const double* vx0, const double* vx1,
                                const double* vy0, const double* vy1) {
  double s1[3], s2[3], w1[3], w2[3], xy[3], yx[3], s1x2[3], s2x2[3];
  for (int d = 0; d < 3; ++d) {
    s1[d] = x1[d] - x0[d];
    s2[d] = y1[d] - y0[d];
    w1[d] = vx1[d] - vx0[d];
    w2[d] = vy1[d] - vy0[d];
    xy[d] = x0[d] - y0[d];
    yx[d] = y0[d] - x0[d];
    s1x2[d] = 2.0 * s1[d];
    s2x2[d] = 2.0 * s2[d];
  }
  const double A = dot3(s1, s1);
  const double B = dot3(s1x2, xy);
  const double C = dot3(s1x2, s2);
  const double D = dot3(s2x2, yx);
  const double E = dot3(s2, s2);
  const double G = C * C - 4.0 * A * E;
  double L1, L2;
  if (A < kEps) {
    L1 = 0.0;
    L2 = E < kEps ? 0.0 : -D / E / 2.0;
  } else if (E < kEps) {
    L2 = 0.0;
    L1 = A < kEps ? 0.0 : -B / A / 2.0;
  } else if (fabs(G) < kEps) {
    double d[4];
    const double* ends[4][2] = {{y0, x0}, {y1, x0}, {y0, x1}, {y1, x1}};
    for (int q = 0; q < 4; ++q) {
      double r[3];
      for (int c = 0; c < 3; ++c) r[c] = ends[q][0][c] - ends[q][1][c];
      d[q] = dot3(r, r);
    }
    int idx = 1;  // first minimum, 1-based
    for (int q = 1; q < 4; ++q)
      if (d[q] < d[idx - 1]) idx = q + 1;
    L1 = floor(idx / 2.0);
    L2 = (double)((idx - 1) % 2);
  } else {
    L2 = (2.0 * A * D + B * C) / G;
    L1 = (C * L2 - B) / A / 2.0;
  }
  L1 = L1 < 0.0 ? 0.0 : (L1 > 1.0 ? 1.0 : L1);
  L2 = L2 < 0.0 ? 0.0 : (L2 > 1.0 ? 1.0 : L2);
  double r[3], vr[3];
  for (int d = 0; d < 3; ++d) {
    r[d] = ((x0[d] + s1[d] * L1) - y0[d]) - s2[d] * L2;
    vr[d] = ((vx0[d] + w1[d] * L1) - vy0[d]) - w2[d] * L2;
  }
  return MinDist{dot3(r, r), 2.0 * dot3(vr, r), L1, L2};
}

__device__ inline bool mobile(int32_t l) { return l == 1 || l == 3; }

// The skip list as sorted keys min(a, b) * 2^32 + max(a, b) of its link
// pairs (built on the host); membership by binary search.
__device__ bool skipped(const int64_t* keys, int64_t nSkip, int64_t a, int64_t b) {
  const int64_t key = a < b ? (a << 32) + b : (b << 32) + a;
  int64_t lo = 0, hi = nSkip;
  while (lo < hi) {
    const int64_t mid = (lo + hi) / 2;
    if (keys[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo < nSkip && keys[lo] == key;
}

// True if candidate (key, pa, pb) beats the current best (bestKey, qa, qb;
// qa < 0: none yet): a larger key, or the same key earlier in loop order.
// Keys are 1/distSq of candidates, so positive (or +Inf), never NaN.
__device__ inline bool better(double key, int64_t pa, int64_t pb, double bestKey, int64_t qa,
                              int64_t qb) {
  if (qa < 0) return true;
  if (key != bestKey) return key > bestKey;
  return pa < qa || (pa == qa && pb < qb);
}

// ---- hinges ---------------------------------------------------------------
// Per node (1-based i = n + 1): its best hinge and its last evaluated hinge.
// Position of hinge (i, j, k): (i, j * maxC + k).
__global__ void k_collision_hinge(int64_t numNode, const int64_t* __restrict__ conn, int64_t ldConn,
                                  const int64_t* __restrict__ links,
                                  const double* __restrict__ coord,
                                  const double* __restrict__ vel,
                                  const int32_t* __restrict__ label, const int64_t* __restrict__ skip,
                                  int64_t nSkip, double collisionDistSq,
                                  CollisionCand* __restrict__ best, CollisionLast* __restrict__ last) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (n >= numNode) return;
  CollisionCand b{0.0, -1, -1, {0, 0, 0, 0, 0, 0}};
  CollisionLast lv{-1, -1, 0.0, 0.0};
  const int64_t i = n + 1;
  if (mobile(label[n])) {
    const int64_t* cn = conn + ldConn * n;
    const int64_t maxC = (ldConn - 1) / 2;
    const double* xi = coord + 3 * n;
    const double* vi = vel + 3 * n;
    for (int64_t j = 1; j <= cn[0]; ++j) {
      // Link 1 starts as connection j and keeps the longer link of every
      // swap below across the k loop (the reference assigns the j loop's
      // link1/connectedNode1/tVecN1 in place, so each k compares against
      // the longest of connections j..k-1).
      int64_t lk1 = cn[2 * j - 1];
      int64_t c1 = links[2 * (lk1 - 1) + (2 - cn[2 * j])];
      double t1[3];
      for (int d = 0; d < 3; ++d) t1[d] = xi[d] - coord[3 * (c1 - 1) + d];
      double tn1 = dot3(t1, t1);
      for (int64_t k = j + 1; k <= cn[0]; ++k) {
        int64_t lk2 = cn[2 * k - 1];
        int64_t c2 = links[2 * (lk2 - 1) + (2 - cn[2 * k])];
        double t2[3];
        for (int d = 0; d < 3; ++d) t2[d] = xi[d] - coord[3 * (c2 - 1) + d];
        double tn2 = dot3(t2, t2);
        if (tn1 < tn2) {  // collapse the shorter one into a point
          const int64_t tl = lk1, tc = c1;
          const double tt = tn1;
          lk1 = lk2;
          c1 = c2;
          tn1 = tn2;
          lk2 = tl;
          c2 = tc;
          tn2 = tt;
        }
        if (skipped(skip, nSkip, lk1, lk2)) continue;
        const double* y0 = coord + 3 * (c2 - 1);
        const double* vy0 = vel + 3 * (c2 - 1);
        const MinDist m = min_distance(xi, coord + 3 * (c1 - 1), y0, y0, vi, vel + 3 * (c1 - 1),
                                       vy0, vy0);
        const int64_t pos = j * maxC + k;
        lv = CollisionLast{i, pos, m.L1, m.L2};
        if (m.distSq < collisionDistSq && m.dDistSqDt < -kEps) {
          const double key = 1.0 / m.distSq;
          if (better(key, i, pos, b.key, b.a, b.b)) {
            b.key = key;
            b.a = i;
            b.b = pos;
            b.info[0] = i;
            b.info[1] = c1;
            b.info[2] = c2;
            b.info[3] = c2;
            b.info[4] = lk1;
            b.info[5] = lk2;
          }
        }
      }
    }
  }
  best[n] = b;
  last[n] = lv;
}

// ---- unconnected pairs ----------------------------------------------------
enum SegF { G_X0 = 0, G_X1 = 3, G_V0 = 6, G_V1 = 9, G_B = 12, G_MID = 15, G_HALF = 18,
            G_COUNT = 19 };

// Pack each link's end points, velocities and b (SoA, ld) and its node
// numbers; flags: bit 0 both ends mobile, bit 1 a skip list is present.
__global__ void k_collision_pack(int64_t numSeg, int64_t ld, const int64_t* __restrict__ links,
                                 const double* __restrict__ coord, const double* __restrict__ vel,
                                 const double* __restrict__ bVec, const int32_t* __restrict__ label,
                                 const int64_t* __restrict__ skip, int64_t nSkip,
                                 double* __restrict__ seg, int64_t* __restrict__ nodes,
                                 int* __restrict__ flags) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= ld) return;
  if (i >= numSeg) {
    for (int f = 0; f < G_COUNT; ++f) seg[f * ld + i] = 0.0;
    nodes[2 * i] = nodes[2 * i + 1] = -1;
    flags[i] = 0;
    return;
  }
  const int64_t a = links[2 * i], b = links[2 * i + 1];
  for (int d = 0; d < 3; ++d) {
    seg[(G_X0 + d) * ld + i] = coord[3 * (a - 1) + d];
    seg[(G_X1 + d) * ld + i] = coord[3 * (b - 1) + d];
    seg[(G_V0 + d) * ld + i] = vel[3 * (a - 1) + d];
    seg[(G_V1 + d) * ld + i] = vel[3 * (b - 1) + d];
    seg[(G_B + d) * ld + i] = bVec[3 * i + d];
  }
  // bounding sphere: midpoint and half length
  double hs = 0.0;
  for (int d = 0; d < 3; ++d) {
    const double pa = coord[3 * (a - 1) + d], pb = coord[3 * (b - 1) + d];
    seg[(G_MID + d) * ld + i] = 0.5 * (pa + pb);
    hs += (pb - pa) * (pb - pa);
  }
  seg[G_HALF * ld + i] = 0.5 * sqrt(hs);
  nodes[2 * i] = a;
  nodes[2 * i + 1] = b;
  int f = (mobile(label[a - 1]) && mobile(label[b - 1])) ? 1 : 0;
  if (nSkip > 0) f |= 2;  // checked per pair by binary search
  flags[i] = f;
}

constexpr int kCT = 64;  // tile side, one wavefront

__device__ __forceinline__ uint32_t spread10(uint32_t x) {
  x &= 1023u;
  x = (x | (x << 16)) & 0x030000FFu;
  x = (x | (x << 8)) & 0x0300F00Fu;
  x = (x | (x << 4)) & 0x030C30C3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

// Morton key of each link's midpoint on a 1024^3 grid over [lo, lo + 1/scale)
// (clamped), and the identity permutation to sort with it.
__global__ void k_collision_keys(const double* __restrict__ seg, int64_t ld, int n, double lox,
                                 double loy, double loz, double scale, uint32_t* __restrict__ keys,
                                 int* __restrict__ idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double lo[3] = {lox, loy, loz};
  uint32_t q[3];
  for (int d = 0; d < 3; ++d) {
    const double v = (seg[(G_MID + d) * ld + i] - lo[d]) * scale;
    q[d] = v <= 0.0 ? 0u : (v >= 1023.0 ? 1023u : (uint32_t)v);
  }
  keys[i] = spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2);
  idx[i] = i;
}

// Bounding box of the end points of the links of each sorted block of 64
// (lo > hi for an empty block): box[6 B + d] low, box[6 B + 3 + d] high.
__global__ void k_collision_boxes(const double* __restrict__ seg, int64_t ld, int n,
                                  const int* __restrict__ perm, int nb, double* __restrict__ box) {
  const int B = blockIdx.x * blockDim.x + threadIdx.x;
  if (B >= nb) return;
  double lo[3] = {1e308, 1e308, 1e308}, hi[3] = {-1e308, -1e308, -1e308};
  const int p1 = min(n, kCT * (B + 1));
  for (int p = kCT * B; p < p1; ++p) {
    const int i = perm[p];
    for (int d = 0; d < 3; ++d) {
      const double a = seg[(G_X0 + d) * ld + i], b = seg[(G_X1 + d) * ld + i];
      lo[d] = fmin(lo[d], fmin(a, b));
      hi[d] = fmax(hi[d], fmax(a, b));
    }
  }
  for (int d = 0; d < 3; ++d) {
    box[6 * B + d] = lo[d];
    box[6 * B + 3 + d] = hi[d];
  }
}

// The reference's filters of pair (i, j), i < j (0-based): links sharing no
// node, both with mobile ends, not a superdislocation
// (DislocationTopologyCollision.jl:176-184), not in the skip list.
__device__ __forceinline__ bool pair_evaluated(const int64_t* ni, const int64_t* nj, int fi, int fj,
                                               const double* sx, const double* sy,
                                               const double* bi, const double* bj,
                                               const int64_t* skip, int64_t nSkip, int64_t i,
                                               int64_t j) {
  if (ni[0] == nj[0] || ni[0] == nj[1] || ni[1] == nj[0] || ni[1] == nj[1]) return false;
  if (!(fi & 1) || !(fj & 1)) return false;
  double dm[3], dp[3];
  for (int d = 0; d < 3; ++d) {
    dm[d] = bi[d] - bj[d];
    dp[d] = bi[d] + bj[d];
  }
  const double par = dot3(sx, sy);
  if ((dot3(dm, dm) < kEps && par > kEps) || (dot3(dp, dp) < kEps && par < kEps)) return false;
  if ((fi & 2) && (fj & 2) && skipped(skip, nSkip, i + 1, j + 1)) return false;
  return true;
}

// Persistent blocks of one wavefront walk the tiles (I, J >= I) of the
// sorted blocks; a tile whose block boxes are farther apart than the
// collision distance (less a rounding margin) is skipped as a whole.  Lane l
// keeps sorted row 64 I + l and pairs it with the 64 columns of block J
// staged in LDS; each lane keeps its best candidate; the wave reduces them
// and writes one partial per block.
__global__ __launch_bounds__(kCT) void k_collision_pairs(
    const double* __restrict__ seg, int64_t ld, int64_t numSeg, const int64_t* __restrict__ nodes,
    const int* __restrict__ flags, const int64_t* __restrict__ skip, int64_t nSkip,
    double collisionDistSq, const int* __restrict__ perm, const double* __restrict__ box,
    int64_t nTiles, int64_t nb, CollisionCand* __restrict__ bestOut) {
  __shared__ double sc[G_COUNT][kCT];
  __shared__ int64_t sn[2][kCT];
  __shared__ int sf[kCT];
  __shared__ int so[kCT];
  __shared__ CollisionCand rb[kCT];
  const int l = threadIdx.x;
  CollisionCand best{0.0, -1, -1, {0, 0, 0, 0, 0, 0}};
  for (int64_t t = blockIdx.x; t < nTiles; t += gridDim.x) {
    // tile t -> (I, J), J >= I, row-major over the upper triangle
    int64_t I = (int64_t)((2 * nb + 1 - sqrt((double)(2 * nb + 1) * (2 * nb + 1) - 8.0 * t)) / 2);
    while (I > 0 && I * nb - I * (I - 1) / 2 > t) --I;
    while ((I + 1) * nb - (I + 1) * I / 2 <= t) ++I;
    const int64_t J = I + (t - (I * nb - I * (I - 1) / 2));
    // Box gap: a lower bound of every point-pair distance of the two blocks,
    // so of every pair's distSq; the margin covers its rounding.
    {
      const double* bI = box + 6 * I;
      const double* bJ = box + 6 * J;
      double g2 = 0.0, mag = 1.0;
      for (int d = 0; d < 3; ++d) {
        const double g = fmax(0.0, fmax(bJ[d] - bI[3 + d], bI[d] - bJ[3 + d]));
        g2 += g * g;
        mag += fmax(fmax(fabs(bI[d]), fabs(bI[3 + d])), fmax(fabs(bJ[d]), fabs(bJ[3 + d])));
      }
      const double gap = sqrt(g2) - 1e-7 * mag;
      if (gap > 0.0 && gap * gap > collisionDistSq) continue;  // uniform over the wave
    }
    __syncthreads();
    {
      const int64_t q = J * kCT + l;
      const int o = q < numSeg ? perm[q] : -1;
      so[l] = o;
      if (o >= 0) {
        for (int f = 0; f < G_COUNT; ++f) sc[f][l] = seg[f * ld + o];
        sn[0][l] = nodes[2 * (int64_t)o];
        sn[1][l] = nodes[2 * (int64_t)o + 1];
        sf[l] = flags[o];
      }
    }
    __syncthreads();
    const int64_t p = I * kCT + l;
    if (p >= numSeg) continue;
    const int64_t i = perm[p];
    const int fi = flags[i];
    if (!(fi & 1)) continue;
    double x0[3], x1[3], v0[3], v1[3], bi[3], sx[3];
    for (int d = 0; d < 3; ++d) {
      x0[d] = seg[(G_X0 + d) * ld + i];
      x1[d] = seg[(G_X1 + d) * ld + i];
      v0[d] = seg[(G_V0 + d) * ld + i];
      v1[d] = seg[(G_V1 + d) * ld + i];
      bi[d] = seg[(G_B + d) * ld + i];
      sx[d] = x1[d] - x0[d];
    }
    const int64_t ni[2] = {nodes[2 * i], nodes[2 * i + 1]};
    const double mi[3] = {seg[G_MID * ld + i], seg[(G_MID + 1) * ld + i],
                          seg[(G_MID + 2) * ld + i]};
    const double hi = seg[G_HALF * ld + i];
    const double mia = fabs(mi[0]) + fabs(mi[1]) + fabs(mi[2]);
    for (int c = (I == J) ? l + 1 : 0; c < kCT; ++c) {
      const int64_t j = so[c];
      if (j < 0) continue;
      double y0[3], y1[3], w0[3], w1[3], bj[3], sy[3];
      for (int d = 0; d < 3; ++d) {
        y0[d] = sc[G_X0 + d][c];
        y1[d] = sc[G_X1 + d][c];
        w0[d] = sc[G_V0 + d][c];
        w1[d] = sc[G_V1 + d][c];
        bj[d] = sc[G_B + d][c];
        sy[d] = y1[d] - y0[d];
      }
      const int64_t nj[2] = {sn[0][c], sn[1][c]};
      const int64_t lo = i < j ? i : j, hiIdx = i < j ? j : i;
      if (!pair_evaluated(ni, nj, fi, sf[c], sx, sy, bi, bj, skip, nSkip, lo, hiIdx)) continue;
      // Bounding spheres: every point pair of the two segments is at least
      // |mi - mj| - hi - hj apart, and the reference's distSq is the squared
      // distance of one such pair, so a pair beyond collisionDist (with a
      // margin for the rounding of the bound) cannot be a candidate.
      {
        double dmv[3];
        for (int d = 0; d < 3; ++d) dmv[d] = mi[d] - sc[G_MID + d][c];
        const double gap = sqrt(dot3(dmv, dmv)) - hi - sc[G_HALF][c];
        const double mj = fabs(sc[G_MID][c]) + fabs(sc[G_MID + 1][c]) + fabs(sc[G_MID + 2][c]);
        const double slack = 1e-7 * (mia + mj + hi + sc[G_HALF][c] + 1.0);
        if (gap - slack > 0.0 && (gap - slack) * (gap - slack) > collisionDistSq) continue;
      }
      // the reference's orientation: the lower link index is segment 1
      const MinDist m = i < j ? min_distance(x0, x1, y0, y1, v0, v1, w0, w1)
                              : min_distance(y0, y1, x0, x1, w0, w1, v0, v1);
      if (m.distSq < collisionDistSq && m.dDistSqDt < -kEps) {
        const double key = 1.0 / m.distSq;
        if (better(key, lo + 1, hiIdx + 1, best.key, best.a, best.b)) {
          best.key = key;
          best.a = lo + 1;
          best.b = hiIdx + 1;
          const int64_t* n1 = i < j ? ni : nj;
          const int64_t* n2 = i < j ? nj : ni;
          best.info[0] = n1[0];
          best.info[1] = n1[1];
          best.info[2] = n2[0];
          best.info[3] = n2[1];
          best.info[4] = lo + 1;
          best.info[5] = hiIdx + 1;
        }
      }
    }
  }
  rb[l] = best;
  __syncthreads();
  if (l == 0) {
    CollisionCand b = rb[0];
    for (int q = 1; q < kCT; ++q) {
      const CollisionCand& o = rb[q];
      if (o.a >= 0 && better(o.key, o.a, o.b, b.key, b.a, b.b)) b = o;
    }
    bestOut[blockIdx.x] = b;
  }
}

constexpr int kLastBlock = 256;

// The last pair the reference evaluates: per row i (one thread, original
// order) the largest j > i that passes the filters, scanned from the end
// over the links with both ends mobile only (prevMobile[j]: the largest
// such link <= j, -1 if none), so fixed links at the end of the array cost
// nothing; the block's largest (i, j) goes to lastOut[blockIdx.x].
__global__ __launch_bounds__(kLastBlock) void k_collision_last(
    const double* __restrict__ seg, int64_t ld, int64_t numSeg, const int64_t* __restrict__ nodes,
    const int* __restrict__ flags, const int* __restrict__ prevMobile,
    const int64_t* __restrict__ skip, int64_t nSkip, CollisionLast* __restrict__ lastOut) {
  __shared__ CollisionLast rl[kLastBlock];
  const int64_t i = blockIdx.x * (int64_t)kLastBlock + threadIdx.x;
  CollisionLast v{-1, -1, 0.0, 0.0};
  if (i < numSeg && (flags[i] & 1)) {
    double bi[3], sx[3];
    for (int d = 0; d < 3; ++d) {
      bi[d] = seg[(G_B + d) * ld + i];
      sx[d] = seg[(G_X1 + d) * ld + i] - seg[(G_X0 + d) * ld + i];
    }
    const int64_t ni[2] = {nodes[2 * i], nodes[2 * i + 1]};
    const int fi = flags[i];
    for (int64_t j = prevMobile[numSeg - 1]; j > i; j = prevMobile[j - 1]) {
      double bj[3], sy[3];
      for (int d = 0; d < 3; ++d) {
        bj[d] = seg[(G_B + d) * ld + j];
        sy[d] = seg[(G_X1 + d) * ld + j] - seg[(G_X0 + d) * ld + j];
      }
      const int64_t nj[2] = {nodes[2 * j], nodes[2 * j + 1]};
      if (pair_evaluated(ni, nj, fi, flags[j], sx, sy, bi, bj, skip, nSkip, i, j)) {
        v = CollisionLast{i + 1, j + 1, 0.0, 0.0};
        break;
      }
    }
  }
  rl[threadIdx.x] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < kLastBlock; ++q) {
      const CollisionLast& w = rl[q];
      if (w.a > v.a || (w.a == v.a && w.b > v.b)) v = w;
    }
    lastOut[blockIdx.x] = v;
  }
}

constexpr int kFinal = 256;

// One block: the best candidate of nBest partials and the last evaluated
// entry of nLast partials, thread-strided then over the threads in order.  With seg (the packed
// links of the pair scan), the last pair's L1, L2 are computed here.
__global__ __launch_bounds__(kFinal) void k_collision_final(
    const CollisionCand* __restrict__ best, int64_t nBest, const CollisionLast* __restrict__ last,
    int64_t nLast, const double* __restrict__ seg, int64_t ld, CollisionCand* __restrict__ ob,
    CollisionLast* __restrict__ ol) {
  __shared__ CollisionCand rb[kFinal];
  __shared__ CollisionLast rl[kFinal];
  const int l = threadIdx.x;
  CollisionCand b{0.0, -1, -1, {0, 0, 0, 0, 0, 0}};
  CollisionLast v{-1, -1, 0.0, 0.0};
  for (int64_t q = l; q < nBest; q += kFinal) {
    const CollisionCand& o = best[q];
    if (o.a >= 0 && better(o.key, o.a, o.b, b.key, b.a, b.b)) b = o;
  }
  for (int64_t q = l; q < nLast; q += kFinal) {
    const CollisionLast& w = last[q];
    if (w.a > v.a || (w.a == v.a && w.b > v.b)) v = w;
  }
  rb[l] = b;
  rl[l] = v;
  __syncthreads();
  if (l == 0) {
    for (int q = 1; q < kFinal; ++q) {
      const CollisionCand& o = rb[q];
      if (o.a >= 0 && better(o.key, o.a, o.b, b.key, b.a, b.b)) b = o;
      const CollisionLast& w = rl[q];
      if (w.a > v.a || (w.a == v.a && w.b > v.b)) v = w;
    }
    if (seg && v.a >= 0) {
      const int64_t i = v.a - 1, j = v.b - 1;
      double p[8][3];
      const int f[8] = {G_X0, G_X1, G_X0, G_X1, G_V0, G_V1, G_V0, G_V1};
      for (int e = 0; e < 8; ++e)
        for (int d = 0; d < 3; ++d) p[e][d] = seg[(f[e] + d) * ld + ((e & 2) ? j : i)];
      const MinDist m = min_distance(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7]);
      v.L1 = m.L1;
      v.L2 = m.L2;
    }
    ob[0] = b;
    ol[0] = v;
  }
}

}  // namespace

void launch_collision_hinge(int64_t numNode, const int64_t* conn, int64_t ldConn,
                            const int64_t* links, const double* coord, const double* vel,
                            const int32_t* label, const int64_t* skip, int64_t nSkip,
                            double collisionDistSq, CollisionCand* best, CollisionLast* last,
                            hipStream_t st) {
  if (numNode <= 0) return;
  hipLaunchKernelGGL(k_collision_hinge, dim3((unsigned)((numNode + 255) / 256)), dim3(256), 0, st,
                     numNode, conn, ldConn, links, coord, vel, label, skip, nSkip,
                     collisionDistSq, best, last);
}

void launch_collision_pack(int64_t numSeg, int64_t ld, const int64_t* links, const double* coord,
                           const double* vel, const double* bVec, const int32_t* label,
                           const int64_t* skip, int64_t nSkip, double* seg, int64_t* nodes,
                           int* flags, hipStream_t st) {
  if (ld <= 0) return;
  hipLaunchKernelGGL(k_collision_pack, dim3((unsigned)((ld + 255) / 256)), dim3(256), 0, st,
                     numSeg, ld, links, coord, vel, bVec, label, skip, nSkip, seg, nodes, flags);
}

int collision_pair_blocks(int64_t numSeg) {
  const int64_t nb = (numSeg + kCT - 1) / kCT;
  const int64_t tiles = nb * (nb + 1) / 2;
  return (int)(tiles < 8192 ? (tiles > 0 ? tiles : 1) : 8192);
}

int collision_last_blocks(int64_t numSeg) {
  return (int)std::max<int64_t>(1, (numSeg + kLastBlock - 1) / kLastBlock);
}

hipError_t collision_sort_bytes(int numSeg, size_t* bytes) {
  *bytes = 0;
  return hipcub::DeviceRadixSort::SortPairs(nullptr, *bytes, (const uint32_t*)nullptr,
                                            (uint32_t*)nullptr, (const int*)nullptr, (int*)nullptr,
                                            numSeg, 0, 30, (hipStream_t)0);
}

hipError_t launch_collision_order(const double* seg, int64_t ld, int numSeg, const double lo[3],
                                  double scale, uint32_t* keys, uint32_t* keysOut, int* idx,
                                  int* perm, void* work, size_t workBytes, double* box,
                                  hipStream_t st) {
  if (numSeg <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_collision_keys, dim3((unsigned)((numSeg + 255) / 256)), dim3(256), 0, st, seg,
                     ld, numSeg, lo[0], lo[1], lo[2], scale, keys, idx);
  size_t bytes = workBytes;
  const hipError_t e =
      hipcub::DeviceRadixSort::SortPairs(work, bytes, keys, keysOut, idx, perm, numSeg, 0, 30, st);
  if (e != hipSuccess) return e;
  const int nb = (numSeg + kCT - 1) / kCT;
  hipLaunchKernelGGL(k_collision_boxes, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, seg, ld,
                     numSeg, perm, nb, box);
  return hipGetLastError();
}

void launch_collision_pairs(const double* seg, int64_t ld, int64_t numSeg, const int64_t* nodes,
                            const int* flags, const int64_t* skip, int64_t nSkip,
                            double collisionDistSq, const int* perm, const double* box, int blocks,
                            CollisionCand* best, hipStream_t st) {
  const int64_t nb = (numSeg + kCT - 1) / kCT;
  const int64_t tiles = nb * (nb + 1) / 2;
  hipLaunchKernelGGL(k_collision_pairs, dim3(blocks), dim3(kCT), 0, st, seg, ld, numSeg, nodes,
                     flags, skip, nSkip, collisionDistSq, perm, box, tiles, nb, best);
}

void launch_collision_last(const double* seg, int64_t ld, int64_t numSeg, const int64_t* nodes,
                           const int* flags, const int* prevMobile, const int64_t* skip,
                           int64_t nSkip, int blocks, CollisionLast* last, hipStream_t st) {
  hipLaunchKernelGGL(k_collision_last, dim3(blocks), dim3(kLastBlock), 0, st, seg, ld, numSeg, nodes,
                     flags, prevMobile, skip, nSkip, last);
}

void launch_collision_final(const CollisionCand* best, int64_t nBest, const CollisionLast* last,
                            int64_t nLast, const double* seg, int64_t ld, CollisionCand* ob,
                            CollisionLast* ol, hipStream_t st) {
  hipLaunchKernelGGL(k_collision_final, dim3(1), dim3(kFinal), 0, st, best, nBest, last, nLast, seg,
                     ld, ob, ol);
}

}  // namespace ddd
