// Internal launch interface between the C-ABI layer and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pair_math.hpp"

namespace ddd {

constexpr int kTile = 64;  // segments per tile side (one wavefront)

// Per-segment SoA record of the pair kernels: field f of segment i at
// seg[f * ld + i].  Fields 0-8 are filled by the host (n1, n2, b), 9-12 by
// prep_segments (unit tangent, 1/L), 13 is 1.0 for an external segment.
enum SegField {
  F_N1X = 0, F_N1Y, F_N1Z, F_N2X, F_N2Y, F_N2Z, F_BX, F_BY, F_BZ,
  F_TX, F_TY, F_TZ, F_TN, F_EXT, F_COUNT
};

// Error bits reported back through the context.
enum : unsigned {
  ERRBIT_NESTED_PARALLEL = 1u,  // a corrective sub-pair was itself near-parallel
};

__host__ __device__ inline int64_t tiles_for(int64_t nb) { return nb * (nb + 1) / 2; }

void launch_prep_segments(double* seg, int64_t ld, int64_t n, hipStream_t st);
// Mixed-precision rule: pairs farther than farDist (midpoint distance) whose
// tangents satisfy |t1 x t2| > sinMin go to the FP32 far-pair quadrature.
// sinMin < 0: no angle condition (near-parallel pairs beyond farDist go to
// the quadrature too, which their parallelism does not affect).
struct MixedRule {
  double farDist, sinMin;
  SigK<float> K;
  double farDist2() const { return farDist * farDist; }
  double sinMin2() const { return sinMin < 0 ? -1.0 : sinMin * sinMin; }
};
// mixed == nullptr: plain FP64.  nExact (mixed only) accumulates the number
// of pairs evaluated on the FP64 path.
// The all-pairs triangle is evaluated in chunks: column blocks [J0, J1),
// i.e. the tiles t in [tiles_for(J0), tiles_for(J1)) of the column-block-
// major enumeration, whose partial sums sit at slot t - tiles_for(J0) of the
// partial buffers and are reduced into segForce before the next chunk.
// Partial sums are kept by super-tile: S x S tiles (rows and columns of S
// blocks, enumerated like the tiles), S = 1 for one tile each.  J0 is a
// multiple of S and super-tile u of super-columns [J0 / S, ceil(J1 / S))
// sits at slot u - u0().  The plain FP64 mode takes S = kSuper on large
// networks (kSuperMin super-tiles or more), the mixed mode S = 1.
constexpr int kSuper = 4;
constexpr int64_t kSuperMin = 16384;
struct TileRange {
  int J0, J1;
  int S = 1;
  int64_t t0() const { return (int64_t)J0 * (J0 + 1) / 2; }
  int64_t t1() const { return (int64_t)J1 * (J1 + 1) / 2; }
  int Q0() const { return J0 / S; }
  int Q1() const { return (J1 + S - 1) / S; }
  int64_t u0() const { return tiles_for(Q0()); }
  int64_t u1() const { return tiles_for(Q1()); }
};
// Near-parallel pairs of the all-pairs tiles, listed per tile and evaluated
// before the tile kernel: off[t] .. off[t + 1] is tile t's range (exclusive
// scan of the count pass over all tiles), its entries at position e - base
// of this chunk's list (base = off[t0]): slot = (k << 6) | lane, tile = t,
// res = [12][total] forces (F1, F2 on the row segment, F3, F4 on the column
// segment).
// fix (total entries) and fixCount: the pairs that need the branch's
// corrective sub-pairs, listed by the evaluation's first pass.
struct SpecialList {
  const long long* off;
  unsigned short* slot;
  int* tile;
  double* res;
  int64_t total;
  int64_t base;
  unsigned* flags;
  int64_t* fix;
  unsigned long long* fixCount;
};
// Count pass (cnt[t] for every tile), then, with off = exclusive scan of cnt
// and the chunk's list sized, the listing pass and the pairs' evaluation
// into res for the tiles of range r.
void launch_special_count(const double* seg, int64_t ld, int n, const MixedRule* mixed,
                          unsigned* cnt, hipStream_t st);
void launch_special_list(const double* seg, int64_t ld, int n, const MixedRule* mixed,
                         const SpecialList& sp, const SegK<double>& K, const TileRange& r,
                         hipStream_t st);
// off[0..n] = exclusive scan of cnt[0..n) in 64-bit sums (off[n] = total);
// work holds scan_work_ints(n) long longs.
int64_t scan_work_ints(int n);
void launch_exclusive_scan(const unsigned* cnt, int n, long long* off, long long* work,
                           hipStream_t st);
// offAt[k] = off[tiles_for(k)] for k = 0..nb (the list offsets at column-block
// boundaries, which size the chunks).
void launch_block_offsets(const long long* off, int nb, long long* offAt, hipStream_t st);
// Workgroups of the persistent mixed-precision tile kernel (two per SIMD)
// and the uint16 staging entries each holds in HBM.
constexpr int kMixedBlocks = 2048;
constexpr int kMixedStage = kTile * kTile;
// The pair tiles of range r into rowPart / colPart (slot t - r.t0()).  Plain
// FP64: a ranged pass and a pass over the tiles it left undone, listed in
// redo (capacity: the range's tile count) with their number in *redoCount.
// Mixed: stage holds kMixedBlocks x kMixedStage entries, counters[0] is the
// tile counter (zeroed here), counters[1] accumulates the FP64 pairs.
void launch_segseg_tiles(const double* seg, int64_t ld, int n, const SegK<double>& K,
                         double* rowPart, double* colPart, const SpecialList& sp,
                         const TileRange& r, hipStream_t st, int* redo,
                         unsigned long long* redoCount, const MixedRule* mixed = nullptr,
                         unsigned short* stage = nullptr, unsigned long long* counters = nullptr);
// The mixed-precision kernels' launches (mixed.hip; used by the two above).
void launch_special_pass_mixed(const double* seg, int64_t ld, int n, const MixedRule& mr,
                               unsigned* cnt, const SpecialList* sp, int64_t t0, int64_t t1,
                               hipStream_t st);
void launch_segseg_tiles_mixed(const double* seg, int64_t ld, int n, const SegK<double>& K,
                               double* rowPart, double* colPart, const SpecialList& sp,
                               const TileRange& r, hipStream_t st, const MixedRule& mr,
                               unsigned short* stage, unsigned long long* counters);
// Adds each segment's partials of range r into out (6 x n): its row role in
// tiles (B, J), J in [max(B, J0), J1), then its column role in tiles (I, B),
// I <= B, if B lies in [J0, J1) -- the fixed order of every chunk.
void launch_segseg_reduce(const double* rowPart, const double* colPart, int n, const TileRange& r,
                          double* out, hipStream_t st);
// Rows-only forces added into out (6 x nRows): each row split over nChunk
// column chunks (segseg_rows_chunks), part holds 6 x nChunk x nRows doubles
// (unused with one chunk).
int segseg_rows_chunks(int n, int nRows);
void launch_segseg_rows(const double* seg, int64_t ld, int n, const int64_t* rows, int nRows,
                        const SegK<double>& K, double* out, double* part, int nChunk,
                        unsigned* flags, hipStream_t st);
// The tile kernel's ranged log / atan / sqrt / division on given arguments
// (which 0-3; division takes x[2i] / x[2i+1]), for bitwise tests.
void launch_ranged_selftest(int which, const double* x, int64_t n, double* y, hipStream_t st);

struct SelfK {
  double a, aSq, mu4pi, nuOmNuInv, Ec, omNuInv;
};
void launch_self_force(const double* seg, int64_t ld, const int64_t* rows, int n,
                       const SelfK& K, double* out, hipStream_t st);

struct MeshDev {
  const double* C;       // 36, column-major
  const double* coord;   // 3 x numNode
  const int64_t* conn;   // 8 x numElem, 1-based
  const double* uHat;    // 3 numNode
  int64_t mx, my, mz;
  double w, h, d;
};
// mode 0: out is 3 x n (pk per segment); mode 1: out is 6 x n, pk/2 to both nodes.
void launch_pk_force(const double* seg, int64_t ld, const int64_t* rows, int n, const MeshDev& M,
                     double* out, int mode, hipStream_t st);
// sigmaHat at P points (3 x P): 9 doubles per point, 3 x 3 column-major.
void launch_sigma_hat(const double* x0, int64_t P, const MeshDev& M, double* out, hipStream_t st);

struct SigmaK {
  double a2, mu4pi, mu4pinu, nuMu4pinu, a2mu8pi;
};
// Stress at P points (3 x P) from nseg segments given as a kSigmaFields-field
// SoA (n1, n2, b uploaded; the rest filled by launch_prep_sigma_segments);
// partial holds nChunk x P x 6 doubles.
int sigma_chunks_for(int64_t P, int nseg);
void launch_sigma_tilde(const double* sseg, int64_t sld, int nseg, const double* x0, int64_t P,
                        const SigmaK& K, double* partial, int nChunk, hipStream_t st);
void launch_sigma_reduce(const double* partial, int nChunk, int64_t P, double* sigma,
                         int accumulate, hipStream_t st);
constexpr int kSigmaFields = 33;
void launch_prep_sigma_segments(double* sseg, int64_t ld, int n, const SigmaK& K, hipStream_t st);
// Segments of the reference's arrays, all on the device (segIdx segIdxLd x 3,
// coord 3 x coordCols, bVec 3 x bVecCols, idx: m 1-based indices or NULL for
// 1..m), gathered into fields n1, n2, b of the kSigmaFields SoA sseg (ld); a
// segment whose indices fall outside the arrays gets NaN coordinates.
void launch_gather_sigma_segments(const int64_t* segIdx, int64_t segIdxLd, int64_t numSeg,
                                  const double* coord, int64_t coordCols, const double* bVec,
                                  int64_t bVecCols, const int64_t* idx, int m, double* sseg,
                                  int64_t ld, hipStream_t st);

// Fixed device scratch of the allocation-free entry point (noalloc.hip):
// 64 MiB per device.  noalloc_init looks up the array's device address with
// `device` current (once per device); noalloc_scratch returns it (NULL if
// the lookup failed or never ran).
constexpr int64_t kNoAllocScratchDoubles = (int64_t)1 << 23;
hipError_t noalloc_init(int device);
double* noalloc_scratch(int device);

void launch_node_gather(const int64_t* conn, int64_t ldConn, int64_t numNode,
                        const int64_t* links, const double* coord, const double* segForce,
                        double sigmaPN, double* nodeForce, hipStream_t st);
// BCC mobility of n nodes (idx: 1-based node list on the device, or NULL
// for nodes 1..n); drag = (edge, screw, climb, line); outputs 3 x n.  With
// a label array (nodeTypeDln per node) fixed and external nodes get zero
// velocity, as deriv! does.
void launch_mobility_bcc(const int64_t* conn, int64_t ldConn, const int64_t* links,
                         const double* coord, const double* bVec, const double* segForce,
                         const int64_t* idx, int64_t n, const double drag[4], double sigmaPN,
                         const int32_t* label, double* nodeForce, double* nodeVel,
                         hipStream_t st);
// Time integration (dynamics.hip): segment end points from device
// coordinates (nodes: 2 per segment, 0-based), the forward-Euler trial, and
// the two maxima of the Euler-trapezoid error test (out[0] = max |distance|,
// out[1] = max |error|; part: 2 trapezoid_parts() entries of scratch).
void launch_gather_coords(double* seg, int64_t ld, int64_t n, const int64_t* nodes,
                          const double* coord, hipStream_t st);
void launch_euler(double* coord, const double* init, const double* vel, double dt, int64_t m,
                  hipStream_t st);
int trapezoid_parts();
// Displacement of triangles at points (displacement.hip).  tri: 13 x tld
// SoA (A, B, C, b, kind: 1 = full triangle, 0 = solid-angle term only);
// partial: nChunk x 3P; launch_chunk_sum adds the chunks into out (3P).
int u_tilde_chunks_for(int64_t P, int ntri);
void launch_u_tilde(const double* tri, int64_t tld, int ntri, const double* pts, int64_t P,
                    double omNuInv8pi, double om2NuOmNuInv8pi, double* partial, int nChunk,
                    hipStream_t st);
void launch_chunk_sum(const double* partial, int nChunk, int64_t count, double* out,
                      int accumulate, hipStream_t st);
// Collision detection (collision.hip).  A candidate: key = 1/distSq, loop
// position (a, b), and the reference's (n1s1, n2s1, n1s2, n2s2, s1, s2); a
// last-evaluated entry: loop position and that pair's L1, L2.  a < 0: none.
struct CollisionCand {
  double key;
  int64_t a, b;
  int64_t info[6];
};
struct CollisionLast {
  int64_t a, b;
  double L1, L2;
};
void launch_collision_hinge(int64_t numNode, const int64_t* conn, int64_t ldConn,
                            const int64_t* links, const double* coord, const double* vel,
                            const int32_t* label, const int64_t* skip, int64_t nSkip,
                            double collisionDistSq, CollisionCand* best, CollisionLast* last,
                            hipStream_t st);
void launch_collision_pack(int64_t numSeg, int64_t ld, const int64_t* links, const double* coord,
                           const double* vel, const double* bVec, const int32_t* label,
                           const int64_t* skip, int64_t nSkip, double* seg, int64_t* nodes,
                           int* flags, hipStream_t st);
int collision_pair_blocks(int64_t numSeg);
int collision_last_blocks(int64_t numSeg);
// Pair scan order: Morton keys of the link midpoints on a 1024^3 grid over
// [lo, lo + 1/scale), radix-sorted into perm (keys, keysOut, idx: numSeg
// each; work: collision_sort_bytes), then the bounding box of each block of
// 64 sorted links (box: 6 doubles per block).
hipError_t collision_sort_bytes(int numSeg, size_t* bytes);
hipError_t launch_collision_order(const double* seg, int64_t ld, int numSeg, const double lo[3],
                                  double scale, uint32_t* keys, uint32_t* keysOut, int* idx,
                                  int* perm, void* work, size_t workBytes, double* box,
                                  hipStream_t st);
void launch_collision_pairs(const double* seg, int64_t ld, int64_t numSeg, const int64_t* nodes,
                            const int* flags, const int64_t* skip, int64_t nSkip,
                            double collisionDistSq, const int* perm, const double* box, int blocks,
                            CollisionCand* best, hipStream_t st);
// prevMobile[j]: the largest link <= j (0-based) with both ends mobile, -1
// if none.
void launch_collision_last(const double* seg, int64_t ld, int64_t numSeg, const int64_t* nodes,
                           const int* flags, const int* prevMobile, const int64_t* skip,
                           int64_t nSkip, int blocks, CollisionLast* last, hipStream_t st);
// seg: the packed links of the pair scan (the last pair's L1, L2 are
// computed from it), or NULL when the entries carry them (hinges).
void launch_collision_final(const CollisionCand* best, int64_t nBest, const CollisionLast* last,
                            int64_t nLast, const double* seg, int64_t ld, CollisionCand* ob,
                            CollisionLast* ol, hipStream_t st);
void launch_trapezoid_max(const double* coord, const double* init, const double* vel,
                          const double* vel0, double dt, int64_t m, unsigned long long* part,
                          double* out, hipStream_t st);

}  // namespace ddd
