// C-ABI layer: contexts, argument checks, packing of the reference's Julia
// arrays into device buffers, kernel sequencing and copy-back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ddd_hip.h"
#include "kernels.hpp"
#include "farfield.hpp"
#include "mlp.hpp"

#include <chrono>
#include <cmath>
#include <random>

using namespace ddd;

namespace {

struct Err {
  int code;
  std::string msg;
};

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      throw Err{e_ == hipErrorOutOfMemory ? DDD_ERR_OOM : DDD_ERR_HIP,                \
                std::string(#expr) + ": " + hipGetErrorString(e_)};                   \
  } while (0)

[[noreturn]] void bad(const std::string& m) { throw Err{DDD_ERR_ARG, m}; }

// A device buffer that grows on demand and is freed with its owner (context,
// far field, network); not copyable.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  template <class T> T* as() const { return static_cast<T*>(p); }
  // Grows by half again at least (fewer reallocations as a network grows),
  // but never past cap when need is within it.
  void ensure(size_t need, size_t cap = SIZE_MAX) {
    if (need <= bytes) return;
    const size_t want = std::max(need, std::min(bytes + bytes / 2, cap));
    if (p) HIPCHK(hipFree(p));
    p = nullptr;
    bytes = 0;
    HIPCHK(hipMalloc(&p, want));
    bytes = want;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// A device buffer of one call's temporaries, freed when it leaves scope.
using ScopedBuf = DevBuf;

}  // namespace

struct ddd_farfield {
  GridDev g{};
  int rn = 1;
  int NP = 0;
  int64_t vol = 0, volH = 0;  // FFT grid points, half-spectrum frequencies
  bool haveOperator = false;
  DevBuf What, tw, cellOf, counts, cursor, start, tmp, order, M, Mc, Lc, L, counters;
  // near-field tile list
  DevBuf mrank, rankCell, mcount, mstart, segS, key, perm, box, rowCount, rowStart, colCount,
      colStart, list, tileIdx, colList, rowPart, colPart;
  // near-field kernel staging, and its list of near-parallel pairs (spCap
  // entries), their per-tile offsets and counts and their forces
  DevBuf stage, spList, specOff, specCnt, spRes, spFix, spFixN;
  long long spCap = 1 << 16;
  int64_t nearTiles = 0;
  void release() {
    DevBuf* all[] = {&What, &tw, &cellOf, &counts, &cursor, &start, &tmp, &order, &M, &Mc, &Lc, &L,
                     &counters, &mrank, &rankCell, &mcount, &mstart, &segS, &key, &perm, &box,
                     &rowCount, &rowStart, &colCount, &colStart, &list, &tileIdx, &colList,
                     &rowPart, &colPart, &stage, &spList, &specOff, &specCnt, &spRes, &spFix,
                     &spFixN};
    for (DevBuf* b : all) b->release();
  }
};

// Split-K workspace of the MLP products (floats): 8 partials of the largest
// split output (kHidden x kMlpOutPad).
constexpr int64_t kGemmWork = 8 * (int64_t)kHidden * kMlpOutPad;

struct ddd_mlp {
  DevBuf P, G, Mm, Vv, Wb;
  float aH = 0.f;  // core radius in cell units of the grid it was trained for
  int cubic = 1;   // learns W at canonical offsets, expanded by cubic symmetry
  int64_t step = 0;
  int cap = 0;
  DevBuf X0, Z1, Z2, Z3, H1, H2, H3, Y, dYb, dH, dZb, lossPer, off3, colPart, gemmWork;
  void ensure_batch(int ns) {
    if (ns <= cap) return;
    X0.ensure(sizeof(unsigned short) * ns * kMlpIn);
    for (DevBuf* b : {&Z1, &Z2, &Z3, &dH}) b->ensure(sizeof(float) * ns * kHidden);
    for (DevBuf* b : {&H1, &H2, &H3, &dZb}) b->ensure(sizeof(unsigned short) * ns * kHidden);
    Y.ensure(sizeof(float) * (int64_t)ns * kMlpOutPad);
    dYb.ensure(sizeof(unsigned short) * (int64_t)ns * kMlpOutPad);
    lossPer.ensure(sizeof(float) * ns);
    off3.ensure(sizeof(int) * 3 * ns);
    colPart.ensure(sizeof(float) * mlp_colsum_part_floats(ns, kMlpOutPad));
    gemmWork.ensure(sizeof(float) * kGemmWork);
    cap = ns;
  }
  void release() {
    DevBuf* all[] = {&P, &G, &Mm, &Vv, &Wb, &X0, &Z1, &Z2, &Z3, &H1, &H2, &H3, &Y, &dYb,
                     &dH, &dZb, &lossPer, &off3, &colPart, &gemmWork};
    for (DevBuf* b : all) b->release();
    cap = 0;
  }
};

struct ddd_ctx {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t st = nullptr;
  std::mutex mu;
  std::string err;
  // segment SoA and pair-kernel scratch
  DevBuf seg, force, rows, out, flags, rowPart;
  // mesh
  DevBuf meshC, meshCoord, meshConn, uHat;
  // field-point stress
  DevBuf sseg, pts, sigPart, sigma;
  // gather and mobility
  DevBuf conn, links, coord, nodeF, bvec, nodeV, nodeIdx;
  // displacement triangles, points, chunk partials, result
  DevBuf uTri, uPts, uPart, uOut;
  // collision detection
  DevBuf cConn, cLinks, cCoord, cVel, cB, cLabel, cSkip, cSeg, cNodes, cFlags, cBest, cLast, cRes;
  DevBuf cKeys, cKeysOut, cIdx, cPerm, cSortWork, cBox, cPrev;
  std::vector<double> h;  // host staging
  DevBuf counter;   // mixed path: tile counter, FP64-pair counter
  DevBuf mixStage;  // mixed path: staging of the FP64 pairs per workgroup
  // near-parallel pairs of the all-pairs tiles (count and offsets over every
  // tile, the scan's work, offsets at the column-block boundaries, fix count)
  DevBuf specCnt, specOff, specWork, specAt, specFixN;
  DevBuf redoN;  // tiles the ranged pass left to the general one (count)
  // one chunk of the all-pairs triangle: its tile partials and its
  // near-parallel list, laid out per chunk by ChunkLayout
  DevBuf chunkMem;
  TileRange pending{0, 0};  // the all-pairs chunk whose reduction is still to run
  double* pendRow = nullptr;  // its partial sums in chunkMem
  double* pendCol = nullptr;
  int chunks = 0;           // chunks of the last all-pairs evaluation
  // events around each chunk's pair-tile kernels (plain FP64 and mixed):
  // their summed time is ddd_stats.tileMs
  std::vector<hipEvent_t> tileEv;
  int tileEvUsed = 0;
  int64_t chunkBytes = (int64_t)32 << 30;
  // device-resident dynamics (deriv!, integrate!)
  DevBuf dCoord, dCoord0, dVel, dVel0, dNodeF, dLinks, dConn, dBvec, dLabel, dSegNodes, dRedPart,
      dRed;
  int64_t dynNodes = -1;
  int64_t dynConnLd = 0;
  double dynDrag[4] = {0, 0, 0, 0};
  // device-resident network
  ddd_opts opts{DDD_MODE_FP64, 0.0, 0.0, nullptr};
  int64_t resN = -1;
  int64_t resLd = 0;
  ddd_consts resK{};
  bool resMesh = false;
  MeshDev resM{};
};

namespace {

int guard(ddd_ctx* ctx, const char* fn, auto&& body) {
  if (!ctx) return DDD_ERR_ARG;
  std::lock_guard<std::mutex> lk(ctx->mu);
  try {
    HIPCHK(hipSetDevice(ctx->device));
    body();
    ctx->err.clear();
    return DDD_OK;
  } catch (const Err& e) {
    ctx->err = std::string(fn) + ": " + e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    ctx->err = std::string(fn) + ": host allocation failed";
    return DDD_ERR_OOM;
  } catch (const std::exception& e) {
    ctx->err = std::string(fn) + ": " + e.what();
    return DDD_ERR_INTERNAL;
  } catch (...) {
    ctx->err = std::string(fn) + ": unknown exception";
    return DDD_ERR_INTERNAL;
  }
}

SegK<double> seg_consts(const ddd_consts& k) {
  const double aSq = k.coreRadSq;
  return SegK<double>{aSq, k.mu4pi, k.mu8pi, aSq * k.mu8pi, k.mu4pinu, aSq * k.mu4pinu};
}

SelfK self_consts(const ddd_consts& k) {
  return SelfK{k.coreRad, k.coreRadSq, k.mu4pi, k.nuOmNuInv, k.coreEnergy, k.omNuInv};
}

int64_t pad64(int64_t n) { return ((n + kTile - 1) / kTile) * kTile + kTile; }

void check_net(const ddd_network* net) {
  if (!net) bad("network is NULL");
  if (net->numSeg < 0 || net->numSeg > net->segIdxLd) bad("numSeg outside [0, size(segIdx, 1)]");
  if (net->numSeg > (int64_t)1 << 30) bad("numSeg too large");
  if (net->numSeg > 0 && (!net->segIdx || !net->coord || !net->bVec))
    bad("segIdx, coord and bVec are required");
}

// Pack segments 0..n-1 into the pair kernels' SoA (fields 0-8 and 13) and
// derive the tangents on the device.
// Pack the network's segments (SoA, leading dimension pad64(numSeg)) into
// dst and derive the per-segment invariants; returns the leading dimension.
int64_t pack_segments(ddd_ctx* c, const ddd_network* net, DevBuf& dst) {
  const int64_t n = net->numSeg, ld = pad64(n);
  c->h.assign((size_t)F_COUNT * ld, 0.0);
  double* h = c->h.data();
  const int64_t* si = net->segIdx;
  const int64_t L = net->segIdxLd;
  for (int64_t i = 0; i < n; ++i) {
    const int64_t lk = si[i], a = si[i + L], b = si[i + 2 * L];
    if (lk < 1 || lk > net->bVecCols) bad("segIdx[:,1] outside bVec");
    if (a < 1 || a > net->coordCols || b < 1 || b > net->coordCols) bad("segIdx node outside coord");
    const double* pa = net->coord + 3 * (a - 1);
    const double* pb = net->coord + 3 * (b - 1);
    const double* pv = net->bVec + 3 * (lk - 1);
    h[F_N1X * ld + i] = pa[0]; h[F_N1Y * ld + i] = pa[1]; h[F_N1Z * ld + i] = pa[2];
    h[F_N2X * ld + i] = pb[0]; h[F_N2Y * ld + i] = pb[1]; h[F_N2Z * ld + i] = pb[2];
    h[F_BX * ld + i] = pv[0]; h[F_BY * ld + i] = pv[1]; h[F_BZ * ld + i] = pv[2];
    h[F_EXT * ld + i] = (net->extSeg && net->extSeg[i]) ? 1.0 : 0.0;
  }
  dst.ensure(sizeof(double) * F_COUNT * ld);
  HIPCHK(hipMemcpyAsync(dst.p, h, sizeof(double) * F_COUNT * ld, hipMemcpyHostToDevice, c->st));
  launch_prep_segments(dst.as<double>(), ld, n, c->st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->st));  // the host staging vector is reused
  return ld;
}

// The one-shot entry points and ddd_net_upload share the segment buffer
// with the resident network, so every upload discards the resident state
// (ddd_net_upload sets it up again afterwards).
void upload_segments(ddd_ctx* c, const ddd_network* net) {
  c->resN = -1;
  c->dynNodes = -1;
  c->resLd = pack_segments(c, net, c->seg);
}

// 0-based row list from 1-based idx, validated.
std::vector<int64_t> rows_from(const int64_t* idx, int64_t nIdx, int64_t numSeg) {
  if (nIdx < 0 || (nIdx > 0 && !idx)) bad("idx/nIdx inconsistent");
  if (nIdx > (int64_t)1 << 30) bad("nIdx too large");
  std::vector<int64_t> r((size_t)nIdx);
  for (int64_t k = 0; k < nIdx; ++k) {
    if (idx[k] < 1 || idx[k] > numSeg) bad("idx entry outside 1:numSeg");
    r[(size_t)k] = idx[k] - 1;
  }
  return r;
}

int64_t* upload_rows(ddd_ctx* c, const std::vector<int64_t>& r) {
  if (r.empty()) return nullptr;
  c->rows.ensure(sizeof(int64_t) * r.size());
  HIPCHK(hipMemcpyAsync(c->rows.p, r.data(), sizeof(int64_t) * r.size(), hipMemcpyHostToDevice,
                        c->st));
  return c->rows.as<int64_t>();
}

double* zeroed(ddd_ctx* c, DevBuf& b, int64_t count) {
  b.ensure(sizeof(double) * std::max<int64_t>(count, 1));
  if (count > 0) HIPCHK(hipMemsetAsync(b.p, 0, sizeof(double) * count, c->st));
  return b.as<double>();
}

unsigned* reset_flags(ddd_ctx* c) {
  c->flags.ensure(sizeof(unsigned));
  HIPCHK(hipMemsetAsync(c->flags.p, 0, sizeof(unsigned), c->st));
  return c->flags.as<unsigned>();
}

void check_flags(ddd_ctx* c) {
  unsigned f = 0;
  HIPCHK(hipMemcpyAsync(&f, c->flags.p, sizeof(unsigned), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  if (f & ERRBIT_NESTED_PARALLEL)
    throw Err{DDD_ERR_NUMERIC, "a corrective sub-pair of the near-parallel branch was itself "
                               "near-parallel (results written, unreliable)"};
}

MixedRule mixed_rule(const ddd_consts& k, const ddd_opts& o) {
  MixedRule r;
  r.farDist = o.farDist > 0 ? o.farDist : 5.0 * k.coreRad;
  r.sinMin = o.sinMin != 0 ? o.sinMin : 1e-3;
  // the FP32 quadrature's reciprocals take a^2 and larger arguments
  if (!(k.coreRadSq >= 0x1p-60 && k.coreRadSq <= 0x1p60))
    bad("the FP32 quadrature needs a squared core radius in [2^-60, 2^60]");
  r.K = SigK<float>{float(k.coreRadSq), float(k.mu4pi), float(k.mu4pinu), float(k.nuMu4pinu),
                    float(k.coreRadSq * k.mu8pi)};
  return r;
}

void check_opts(const ddd_opts* o) {
  if (!o) return;
  if (o->mode != DDD_MODE_FP64 && o->mode != DDD_MODE_MIXED && o->mode != DDD_MODE_HYBRID)
    bad("unknown mode");
  if (!(o->farDist >= 0) || std::isnan(o->sinMin)) bad("farDist must be >= 0 and sinMin a number");
  if (o->mode == DDD_MODE_HYBRID) {
    if (!o->farfield) bad("hybrid mode needs a farfield (ddd_farfield_create)");
    if (!o->farfield->haveOperator) bad("the farfield has no operator (ddd_farfield_set_operator)");
  }
}

bool is_hybrid(const ddd_opts* o) { return o && o->mode == DDD_MODE_HYBRID; }

// Hybrid pair stage: cells, the near-field tile list, moments, far-field
// expansions, near-field tiles, their near-parallel pairs and their
// reduction (adds into dev).  The far-field forces are added by
// hybrid_apply.  Two host synchronisations: the number of near tiles sizes
// the partial-sum buffers, the length of the near-parallel list sizes its
// evaluation.
void hybrid_pairs(ddd_ctx* c, const ddd_consts& k, int64_t n, double* dev, unsigned* flags,
                  const ddd_opts& o) {
  ddd_farfield* ff = o.farfield;
  const int nc = ff->g.G * ff->g.G * ff->g.G;
  const int64_t n1 = std::max<int64_t>(n, 1);
  ff->cellOf.ensure(sizeof(int) * n1);
  ff->tmp.ensure(sizeof(int) * n1);
  ff->order.ensure(sizeof(int) * n1);
  ff->counts.ensure(sizeof(unsigned) * nc);
  ff->cursor.ensure(sizeof(unsigned) * nc);
  ff->start.ensure(sizeof(int) * (nc + 1));
  ff->M.ensure(sizeof(float) * kMoments * nc);
  ff->L.ensure(sizeof(float) * kLocal * nc);
  ff->Mc.ensure(sizeof(float2) * kMomPairs * ff->vol);
  ff->Lc.ensure(sizeof(float2) * kLocPairs * ff->vol);
  ff->counters.ensure(4 * sizeof(unsigned long long));
  ff->stage.ensure(sizeof(unsigned short) * (size_t)kNearGrid * kNearStage);
  const int64_t ldS = pad64(n1);
  const int64_t nbS = ldS / kTile;
  ff->mcount.ensure(sizeof(unsigned) * nc);
  ff->mstart.ensure(sizeof(int) * (nc + 1));
  ff->segS.ensure(sizeof(double) * F_COUNT * ldS);
  ff->key.ensure(sizeof(int) * ldS);
  ff->perm.ensure(sizeof(int) * ldS);
  ff->box.ensure(sizeof(int) * 6 * nbS);
  ff->rowCount.ensure(sizeof(unsigned) * nbS);
  ff->rowStart.ensure(sizeof(int) * (nbS + 1));
  ff->colCount.ensure(sizeof(unsigned) * nbS);
  ff->colStart.ensure(sizeof(int) * (nbS + 1));
  ff->tileIdx.ensure(sizeof(int) * nbS * nbS);
  // near pairs, of which FP64, and the near-field kernels' tile counters
  HIPCHK(hipMemsetAsync(ff->counters.p, 0, 4 * sizeof(unsigned long long), c->st));
  const double* seg = c->seg.as<double>();
  launch_cells(seg, c->resLd, (int)n, ff->g, ff->cellOf.as<int>(), ff->counts.as<unsigned>(),
               ff->cursor.as<unsigned>(), ff->start.as<int>(), ff->tmp.as<int>(),
               ff->order.as<int>(), c->st);
  NearListDev nl{};
  nl.cellOf = ff->cellOf.as<int>();
  nl.counts = ff->counts.as<unsigned>();
  nl.start = ff->start.as<int>();
  nl.order = ff->order.as<int>();
  nl.mrank = ff->mrank.as<int>();
  nl.rankCell = ff->rankCell.as<int>();
  nl.mcount = ff->mcount.as<unsigned>();
  nl.mstart = ff->mstart.as<int>();
  nl.ldS = ldS;
  nl.segS = ff->segS.as<double>();
  nl.key = ff->key.as<int>();
  nl.perm = ff->perm.as<int>();
  nl.box = ff->box.as<int>();
  nl.rowCount = ff->rowCount.as<unsigned>();
  nl.rowStart = ff->rowStart.as<int>();
  nl.colCount = ff->colCount.as<unsigned>();
  nl.colStart = ff->colStart.as<int>();
  nl.tileIdx = ff->tileIdx.as<int>();
  // Sorted copy and per-row tile counts; the total sizes the lists.
  launch_near_list_count(seg, c->resLd, (int)n, ff->g, ff->rn, nl, c->st);
  int tiles = 0;
  HIPCHK(hipMemcpyAsync(&tiles, ff->rowStart.as<int>() + nbS, sizeof(int), hipMemcpyDeviceToHost,
                        c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  ff->nearTiles = tiles;
  ff->list.ensure(sizeof(int2) * std::max(tiles, 1));
  ff->colList.ensure(sizeof(int) * std::max(tiles, 1));
  ff->rowPart.ensure(sizeof(double) * 6 * kTile * std::max(tiles, 1));
  ff->colPart.ensure(sizeof(double) * 6 * kTile * std::max(tiles, 1));
  ff->specCnt.ensure(sizeof(int) * std::max(tiles, 1));
  ff->specOff.ensure(sizeof(long long) * std::max(tiles, 1));
  nl.list = ff->list.as<int2>();
  nl.colList = ff->colList.as<int>();
  launch_near_list_fill(nl, c->st);
  launch_moments(seg, c->resLd, ff->g, ff->start.as<int>(), ff->order.as<int>(), ff->M.as<float>(),
                 c->st);
  launch_far_field(ff->M.as<float>(), ff->g, ff->What.as<float2>(), ff->tw.as<float2>(),
                   ff->Mc.as<float2>(), ff->Lc.as<float2>(), ff->L.as<float>(), c->st);
  const MixedRule mr = mixed_rule(k, o);
  const SegK<double> K = seg_consts(k);
  // Near-field tiles; their near-parallel pairs are listed, and the list's
  // length (read back: a second host synchronisation) sizes their
  // evaluation.  A list that outgrew its capacity is enlarged and the tiles
  // run again (a one-off as the network changes).
  long long total = 0;
  ff->spCap = std::max(ff->spCap, 16LL * tiles);  // ~12 per tile on random networks
  for (;;) {
    ff->spList.ensure(sizeof(unsigned long long) * ff->spCap);
    launch_near_tiles(ff->segS.as<double>(), ldS, ff->key.as<int>(), ff->rn, ff->list.as<int2>(),
                      tiles, K, mr, ff->rowPart.as<double>(), ff->colPart.as<double>(),
                      ff->stage.as<unsigned short>(), ff->spList.as<unsigned long long>(), ff->spCap,
                      ff->specOff.as<long long>(), ff->specCnt.as<int>(),
                      ff->counters.as<unsigned long long>(), c->st);
    HIPCHK(hipMemcpyAsync(&total, ff->counters.as<unsigned long long>() + 3, sizeof(total),
                          hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (total <= ff->spCap) break;
    ff->spCap = total + total / 4;
    HIPCHK(hipMemsetAsync(ff->counters.p, 0, 4 * sizeof(unsigned long long), c->st));
  }
  ff->spRes.ensure(sizeof(double) * 12 * std::max<long long>(ff->spCap, 1));
  ff->spFix.ensure(sizeof(long long) * std::max<long long>(ff->spCap, 1));
  ff->spFixN.ensure(sizeof(unsigned long long));
  launch_near_special(ff->segS.as<double>(), ldS, ff->list.as<int2>(), tiles, K,
                      ff->spList.as<unsigned long long>(), total, ff->spCap,
                      ff->specOff.as<long long>(), ff->specCnt.as<int>(), ff->spRes.as<double>(),
                      ff->spFix.as<long long>(), ff->spFixN.as<unsigned long long>(),
                      ff->rowPart.as<double>(), ff->colPart.as<double>(), flags, c->st);
  launch_near_reduce(nl, nc, ff->rowPart.as<double>(), ff->colPart.as<double>(), dev, c->st);
  HIPCHK(hipGetLastError());
}

void hybrid_apply(ddd_ctx* c, int64_t n, double* dev, const ddd_opts& o) {
  ddd_farfield* ff = o.farfield;
  launch_far_apply(c->seg.as<double>(), c->resLd, (int)n, ff->g, ff->cellOf.as<int>(),
                   ff->L.as<float>(), dev, c->st);
  HIPCHK(hipGetLastError());
}

// Device memory of one chunk of the all-pairs triangle, one buffer carved
// per chunk: row and column partial sums (6 doubles x 64 lanes each per
// tile; the plain FP64 mode on large networks keeps them per super-tile,
// kSuper rows and kSuper columns of them, counted here as kSuper tiles),
// the list of tiles or super-tiles
// the ranged pass leaves to the general one (one int each) and the
// near-parallel list (slot, tile, 12 forces and a fix-list entry per pair),
// each part 256-byte aligned.  A chunk takes whole column blocks (whole
// super-columns of kSuper blocks in the plain mode) while its layout stays
// within the context's chunkBytes (default 32 GiB), so the buffer never
// exceeds the budget; a single column above it makes a chunk of its own.
struct ChunkLayout {
  int64_t row = 0, col = 0, redo = 0, slot = 0, tile = 0, res = 0, fix = 0, bytes = 0;
  ChunkLayout(int64_t tiles, int64_t spec) {
    int64_t at = 0;
    auto part = [&](int64_t b) {
      const int64_t o = at;
      at += (b + 255) / 256 * 256;
      return o;
    };
    row = part(sizeof(double) * 6 * kTile * tiles);
    col = part(sizeof(double) * 6 * kTile * tiles);
    redo = part(sizeof(int) * tiles);
    slot = part(sizeof(unsigned short) * spec);
    tile = part(sizeof(int) * spec);
    res = part(sizeof(double) * 12 * spec);
    fix = part(sizeof(int64_t) * spec);
    bytes = at;
  }
};

// Pair tiles of the whole network (mode from opts), chunk by chunk: a count
// pass and scan over every tile for the near-parallel pairs (one host
// synchronisation reads the list offsets at the column-block boundaries,
// which size the chunks), then per chunk the listing and evaluation of its
// near-parallel pairs, the tile kernel (which adds them to its partial
// sums) and the reduction of its partials into dev -- except the last
// chunk's reduction, which segseg_finish runs (the bench times the tile
// stage on its own; with one chunk, as at 131,072 segments, that is the
// whole reduction).  Memory is bounded by chunkBytes instead of growing
// with the N^2 / 8192 tiles.
void segseg_tiles(ddd_ctx* c, const ddd_consts& k, int64_t n, unsigned* flags,
                  const ddd_opts* opts, double* dev) {
  c->pending = TileRange{0, 0};
  if (n < 2) return;
  const int64_t nb = (n + kTile - 1) / kTile;
  const int64_t T = tiles_for(nb);
  if (T + 1 > INT32_MAX) bad("too many segments for the tile list");
  c->counter.ensure(2 * sizeof(unsigned long long));
  c->specCnt.ensure(sizeof(unsigned) * T);
  c->specOff.ensure(sizeof(long long) * (T + 1));
  c->specAt.ensure(sizeof(long long) * (nb + 1));
  const bool mixed = opts && opts->mode == DDD_MODE_MIXED;
  const MixedRule r = mixed ? mixed_rule(k, *opts) : MixedRule{};
  launch_special_count(c->seg.as<double>(), c->resLd, (int)n, mixed ? &r : nullptr,
                       c->specCnt.as<unsigned>(), c->st);
  c->specWork.ensure(sizeof(long long) * scan_work_ints((int)T));
  launch_exclusive_scan(c->specCnt.as<unsigned>(), (int)T, c->specOff.as<long long>(),
                        c->specWork.as<long long>(), c->st);
  launch_block_offsets(c->specOff.as<long long>(), (int)nb, c->specAt.as<long long>(), c->st);
  std::vector<long long> offAt((size_t)nb + 1);
  HIPCHK(hipMemcpyAsync(offAt.data(), c->specAt.p, sizeof(long long) * (nb + 1),
                        hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  // chunks of whole column blocks within the budget; the plain FP64 mode
  // on networks of kSuperMin super-tiles or more keeps its partials by
  // super-tile (kSuper x kSuper tiles, 24 KiB), so its chunks take whole
  // super-columns of kSuper column blocks; smaller networks and the mixed
  // mode keep them by tile
  std::vector<TileRange> chunks;
  const int step = (!mixed && tiles_for((nb + kSuper - 1) / kSuper) >= kSuperMin) ? kSuper : 1;
  auto range = [&](int J0, int J1) { return TileRange{J0, J1, step}; };
  auto layout = [&](int J0, int J1) {
    const TileRange q = range(J0, J1);
    return ChunkLayout((q.u1() - q.u0()) * step, offAt[J1] - offAt[J0]);
  };
  int64_t most = 0;
  for (int J0 = 0; J0 < nb;) {
    int J1 = (int)std::min<int64_t>(J0 + step, nb);
    while (J1 < nb && layout(J0, (int)std::min<int64_t>(J1 + step, nb)).bytes <= c->chunkBytes)
      J1 = (int)std::min<int64_t>(J1 + step, nb);
    chunks.push_back(range(J0, J1));
    most = std::max(most, layout(J0, J1).bytes);
    J0 = J1;
  }
  c->chunks = (int)chunks.size();
  c->chunkMem.ensure((size_t)most, (size_t)std::max(most, c->chunkBytes));
  c->specFixN.ensure(sizeof(unsigned long long));
  c->redoN.ensure(sizeof(unsigned long long));
  if (mixed) {
    HIPCHK(hipMemsetAsync(c->counter.p, 0, 2 * sizeof(unsigned long long), c->st));
    c->mixStage.ensure(sizeof(unsigned short) * kMixedBlocks * kMixedStage);
  }
  while (c->tileEv.size() < 2 * chunks.size()) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    c->tileEv.push_back(e);
  }
  c->tileEvUsed = 0;
  char* mem = c->chunkMem.as<char>();
  for (size_t q = 0; q < chunks.size(); ++q) {
    const TileRange& ch = chunks[q];
    const ChunkLayout L = layout(ch.J0, ch.J1);
    double* rowPart = reinterpret_cast<double*>(mem + L.row);
    double* colPart = reinterpret_cast<double*>(mem + L.col);
    SpecialList sp{c->specOff.as<long long>(), reinterpret_cast<unsigned short*>(mem + L.slot),
                   reinterpret_cast<int*>(mem + L.tile), reinterpret_cast<double*>(mem + L.res),
                   (int64_t)(offAt[ch.J1] - offAt[ch.J0]), (int64_t)offAt[ch.J0], flags,
                   reinterpret_cast<int64_t*>(mem + L.fix), c->specFixN.as<unsigned long long>()};
    launch_special_list(c->seg.as<double>(), c->resLd, (int)n, mixed ? &r : nullptr, sp,
                        seg_consts(k), ch, c->st);
    HIPCHK(hipEventRecord(c->tileEv[2 * q], c->st));
    launch_segseg_tiles(c->seg.as<double>(), c->resLd, (int)n, seg_consts(k), rowPart, colPart, sp,
                        ch, c->st, reinterpret_cast<int*>(mem + L.redo),
                        c->redoN.as<unsigned long long>(), mixed ? &r : nullptr,
                        mixed ? c->mixStage.as<unsigned short>() : nullptr,
                        mixed ? c->counter.as<unsigned long long>() : nullptr);
    HIPCHK(hipEventRecord(c->tileEv[2 * q + 1], c->st));
    c->tileEvUsed = (int)q + 1;
    if (q + 1 < chunks.size()) {
      launch_segseg_reduce(rowPart, colPart, (int)n, ch, dev, c->st);
    } else {
      c->pending = ch;
      c->pendRow = rowPart;
      c->pendCol = colPart;
    }
  }
  HIPCHK(hipGetLastError());
}

// The last chunk's reduction (see segseg_tiles).
void segseg_finish(ddd_ctx* c, int64_t n, double* dev) {
  if (c->pending.J1 > 0)
    launch_segseg_reduce(c->pendRow, c->pendCol, (int)n, c->pending, dev, c->st);
  c->pending = TileRange{0, 0};
  HIPCHK(hipGetLastError());
}

// Whole-network seg-seg into dev (6 x n, accumulated).
void segseg_all(ddd_ctx* c, const ddd_consts& k, int64_t n, double* dev, unsigned* flags,
                const ddd_opts* opts = nullptr) {
  if (n < 2) return;
  if (is_hybrid(opts)) {
    hybrid_pairs(c, k, n, dev, flags, *opts);
    hybrid_apply(c, n, dev, *opts);
    return;
  }
  segseg_tiles(c, k, n, flags, opts, dev);
  segseg_finish(c, n, dev);
}

// Rows-only seg-seg forces of the rows dr (nIdx of them) added into dev.
void launch_rows(ddd_ctx* c, int64_t n, const int64_t* dr, int64_t nIdx, const ddd_consts& k,
                 double* dev, unsigned* fl) {
  const int nChunk = segseg_rows_chunks((int)n, (int)nIdx);
  if (nChunk > 1) c->rowPart.ensure(sizeof(double) * 6 * (size_t)nChunk * (size_t)nIdx);
  launch_segseg_rows(c->seg.as<double>(), c->resLd, (int)n, dr, (int)nIdx, seg_consts(k), dev,
                     nChunk > 1 ? c->rowPart.as<double>() : nullptr, nChunk, fl, c->st);
}

// Copy a 6 x m device result into the caller's 3 x 2 x outCols array.
void store6(ddd_ctx* c, const double* dev, int64_t m, const std::vector<int64_t>* rows,
            double* out, int64_t outCols, int flags) {
  std::vector<double> hv((size_t)(6 * m));
  if (m > 0)
    HIPCHK(hipMemcpyAsync(hv.data(), dev, sizeof(double) * 6 * m, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  const bool acc = flags & DDD_ACCUMULATE;
  const bool byPos = (flags & DDD_OUT_BY_POSITION) || !rows;
  for (int64_t k = 0; k < m; ++k) {
    const int64_t col = byPos ? k : (*rows)[(size_t)k];
    if (col >= outCols) bad("output has too few columns");
    double* o = out + 6 * col;
    for (int q = 0; q < 6; ++q) o[q] = acc ? o[q] + hv[(size_t)(6 * k + q)] : hv[(size_t)(6 * k + q)];
  }
}

void check_out(const double* out, int64_t outCols, int64_t need) {
  if (!out) bad("output is NULL");
  if (outCols < need) bad("output has too few columns");
}

MeshDev upload_mesh(ddd_ctx* c, const ddd_mesh* m) {
  if (!m->C || !m->coord || !m->connectivity || !m->uHat) bad("mesh arrays are required");
  if (m->mx < 1 || m->my < 1 || m->mz < 1 || m->numElem != m->mx * m->my * m->mz)
    bad("mesh element counts inconsistent");
  if (!(m->w > 0 && m->h > 0 && m->d > 0)) bad("mesh element sizes must be positive");
  for (int64_t q = 0; q < 8 * m->numElem; ++q)
    if (m->connectivity[q] < 1 || m->connectivity[q] > m->numNode) bad("mesh connectivity entry");
  c->meshC.ensure(sizeof(double) * 36);
  c->meshCoord.ensure(sizeof(double) * 3 * m->numNode);
  c->meshConn.ensure(sizeof(int64_t) * 8 * m->numElem);
  c->uHat.ensure(sizeof(double) * 3 * m->numNode);
  HIPCHK(hipMemcpyAsync(c->meshC.p, m->C, sizeof(double) * 36, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->meshCoord.p, m->coord, sizeof(double) * 3 * m->numNode,
                        hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->meshConn.p, m->connectivity, sizeof(int64_t) * 8 * m->numElem,
                        hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->uHat.p, m->uHat, sizeof(double) * 3 * m->numNode,
                        hipMemcpyHostToDevice, c->st));
  return MeshDev{c->meshC.as<double>(), c->meshCoord.as<double>(), c->meshConn.as<int64_t>(),
                 c->uHat.as<double>(), m->mx, m->my, m->mz, m->w, m->h, m->d};
}

}  // namespace

extern "C" {

int ddd_version(void) { return DDD_ABI_VERSION; }

int ddd_ctx_create(ddd_ctx** out, int device) {
  if (!out) return DDD_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0 || device < 0 || device >= n)
    return DDD_ERR_NO_DEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return DDD_ERR_NO_DEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return DDD_ERR_NO_DEVICE;
  ddd_ctx* c = new (std::nothrow) ddd_ctx();
  if (!c) return DDD_ERR_OOM;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return DDD_ERR_HIP;
  }
  c->st = c->own;
  // The allocation-free entry point's scratch: its address on this device,
  // looked up once per device; a failure is reported by that entry point.
  (void)noalloc_init(device);
  *out = c;
  return DDD_OK;
}

int ddd_ctx_destroy(ddd_ctx* c) {
  if (!c) return DDD_OK;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->st);
    DevBuf* all[] = {&c->counter, &c->seg, &c->chunkMem, &c->force, &c->rows, &c->out,
                     &c->flags, &c->meshC, &c->meshCoord, &c->meshConn, &c->uHat, &c->sseg,
                     &c->pts, &c->sigPart, &c->sigma, &c->conn, &c->links, &c->coord, &c->nodeF,
                     &c->specCnt, &c->specOff, &c->specWork};
    for (DevBuf* b : all) b->release();
    for (hipEvent_t e : c->tileEv) (void)hipEventDestroy(e);
    if (c->own) (void)hipStreamDestroy(c->own);
  }
  delete c;
  return DDD_OK;
}

int ddd_ctx_set_stream(ddd_ctx* c, void* stream) {
  return guard(c, "ddd_ctx_set_stream", [&] {
    HIPCHK(hipStreamSynchronize(c->st));
    c->st = stream ? static_cast<hipStream_t>(stream) : c->own;
  });
}

void* ddd_ctx_stream(ddd_ctx* c) { return c ? (void*)c->st : nullptr; }

int ddd_ctx_set_chunk_bytes(ddd_ctx* c, int64_t bytes) {
  return guard(c, "ddd_ctx_set_chunk_bytes", [&] {
    if (bytes < 0) bad("bytes < 0");
    c->chunkBytes = bytes > 0 ? bytes : ((int64_t)32 << 30);
  });
}

const char* ddd_last_error(ddd_ctx* c) { return c ? c->err.c_str() : "no context"; }

int ddd_segseg_force(ddd_ctx* c, const ddd_consts* k, const ddd_network* net, const int64_t* idx,
                     int64_t nIdx, double* out, int64_t outCols, int flags) {
  return guard(c, "ddd_segseg_force", [&] {
    if (!k) bad("consts is NULL");
    check_net(net);
    const int64_t n = net->numSeg;
    upload_segments(c, net);
    unsigned* fl = reset_flags(c);
    if (!idx) {
      check_out(out, outCols, n);
      double* dev = zeroed(c, c->force, 6 * n);
      segseg_all(c, *k, n, dev, fl);
      check_flags(c);
      store6(c, dev, n, nullptr, out, outCols, flags);
    } else {
      std::vector<int64_t> r = rows_from(idx, nIdx, n);
      check_out(out, outCols, (flags & DDD_OUT_BY_POSITION) ? nIdx : 0);
      int64_t* dr = upload_rows(c, r);
      double* dev = zeroed(c, c->out, 6 * nIdx);
      launch_rows(c, n, dr, nIdx, *k, dev, fl);
      HIPCHK(hipGetLastError());
      check_flags(c);
      store6(c, dev, nIdx, &r, out, outCols, flags);
    }
    c->resN = -1;
  });
}

int ddd_self_force(ddd_ctx* c, const ddd_consts* k, const ddd_network* net, const int64_t* idx,
                   int64_t nIdx, double* out, int64_t outCols, int flags) {
  return guard(c, "ddd_self_force", [&] {
    if (!k) bad("consts is NULL");
    check_net(net);
    const int64_t n = net->numSeg;
    upload_segments(c, net);
    std::vector<int64_t> r;
    if (idx) r = rows_from(idx, nIdx, n);
    const int64_t m = idx ? nIdx : n;
    check_out(out, outCols, (idx && !(flags & DDD_OUT_BY_POSITION)) ? 0 : m);
    int64_t* dr = upload_rows(c, r);
    double* dev = zeroed(c, c->out, 6 * m);
    launch_self_force(c->seg.as<double>(), c->resLd, idx ? dr : nullptr, (int)m, self_consts(*k),
                      dev, c->st);
    HIPCHK(hipGetLastError());
    store6(c, dev, m, idx ? &r : nullptr, out, outCols, flags);
    c->resN = -1;
  });
}

int ddd_pk_force(ddd_ctx* c, const ddd_network* net, const ddd_mesh* mesh, const int64_t* idx,
                 int64_t nIdx, double* out, int64_t outCols, int flags) {
  return guard(c, "ddd_pk_force", [&] {
    check_net(net);
    if (!mesh) bad("mesh is NULL");
    const int64_t n = net->numSeg;
    upload_segments(c, net);
    MeshDev M = upload_mesh(c, mesh);
    std::vector<int64_t> r;
    if (idx) r = rows_from(idx, nIdx, n);
    const int64_t m = idx ? nIdx : n;
    const bool half = flags & DDD_PK_HALF_TO_NODES;
    const bool byPos = (flags & DDD_OUT_BY_POSITION) || !idx;
    check_out(out, outCols, byPos ? m : 0);
    int64_t* dr = upload_rows(c, r);
    const int w = half ? 6 : 3;
    double* dev = zeroed(c, c->out, w * m);
    launch_pk_force(c->seg.as<double>(), c->resLd, idx ? dr : nullptr, (int)m, M, dev, half ? 1 : 0,
                    c->st);
    HIPCHK(hipGetLastError());
    if (half) {
      store6(c, dev, m, idx ? &r : nullptr, out, outCols, flags);
    } else {
      std::vector<double> hv((size_t)(3 * m));
      if (m > 0)
        HIPCHK(hipMemcpyAsync(hv.data(), dev, sizeof(double) * 3 * m, hipMemcpyDeviceToHost, c->st));
      HIPCHK(hipStreamSynchronize(c->st));
      const bool acc = flags & DDD_ACCUMULATE;
      for (int64_t q = 0; q < m; ++q) {
        const int64_t col = byPos ? q : r[(size_t)q];
        if (col >= outCols) bad("output has too few columns");
        for (int e = 0; e < 3; ++e)
          out[3 * col + e] = acc ? out[3 * col + e] + hv[(size_t)(3 * q + e)] : hv[(size_t)(3 * q + e)];
      }
    }
    c->resN = -1;
  });
}

int ddd_sigma_hat(ddd_ctx* c, const ddd_mesh* mesh, const double* x0, int64_t P, double* sigma) {
  return guard(c, "ddd_sigma_hat", [&] {
    if (!mesh) bad("mesh is NULL");
    if (P < 0 || (P > 0 && (!x0 || !sigma))) bad("points/sigma inconsistent");
    if (P > ((int64_t)1 << 32)) bad("too many points");
    MeshDev M = upload_mesh(c, mesh);
    if (P == 0) {
      HIPCHK(hipStreamSynchronize(c->st));
      return;
    }
    c->pts.ensure(sizeof(double) * 3 * P);
    c->sigma.ensure(sizeof(double) * 9 * P);
    HIPCHK(hipMemcpyAsync(c->pts.p, x0, sizeof(double) * 3 * P, hipMemcpyHostToDevice, c->st));
    launch_sigma_hat(c->pts.as<double>(), P, M, c->sigma.as<double>(), c->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(sigma, c->sigma.p, sizeof(double) * 9 * P, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

int ddd_selftest_ranged(ddd_ctx* c, int which, const double* x, int64_t n, double* y) {
  return guard(c, "ddd_selftest_ranged", [&] {
    if (which < 0 || which > 3) bad("which must be 0 (log), 1 (atan), 2 (sqrt) or 3 (division)");
    if (n < 0 || (n > 0 && (!x || !y))) bad("x/y inconsistent with n");
    if (n > ((int64_t)1 << 28)) bad("n too large");
    if (n == 0) return;
    const int64_t nin = which == 3 ? 2 * n : n;
    ScopedBuf dx, dy;
    dx.ensure(sizeof(double) * nin);
    dy.ensure(sizeof(double) * n);
    HIPCHK(hipMemcpyAsync(dx.p, x, sizeof(double) * nin, hipMemcpyHostToDevice, c->st));
    launch_ranged_selftest(which, dx.as<double>(), n, dy.as<double>(), c->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(y, dy.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

int ddd_seg_force(ddd_ctx* c, const ddd_consts* k, const ddd_network* net, const ddd_mesh* mesh,
                  const int64_t* idx, int64_t nIdx, double* out, int64_t outCols, int flags) {
  return guard(c, "ddd_seg_force", [&] {
    if (!k) bad("consts is NULL");
    check_net(net);
    const int64_t n = net->numSeg;
    upload_segments(c, net);
    MeshDev M{};
    if (mesh) M = upload_mesh(c, mesh);
    unsigned* fl = reset_flags(c);
    if (!idx) {
      check_out(out, outCols, n);
      double* dev = zeroed(c, c->force, 6 * n);
      if (mesh) launch_pk_force(c->seg.as<double>(), c->resLd, nullptr, (int)n, M, dev, 1, c->st);
      launch_self_force(c->seg.as<double>(), c->resLd, nullptr, (int)n, self_consts(*k), dev, c->st);
      HIPCHK(hipGetLastError());
      segseg_all(c, *k, n, dev, fl);
      check_flags(c);
      store6(c, dev, n, nullptr, out, outCols, flags);
    } else {
      std::vector<int64_t> r = rows_from(idx, nIdx, n);
      check_out(out, outCols, (flags & DDD_OUT_BY_POSITION) ? nIdx : 0);
      int64_t* dr = upload_rows(c, r);
      double* dev = zeroed(c, c->out, 6 * nIdx);
      if (mesh) launch_pk_force(c->seg.as<double>(), c->resLd, dr, (int)nIdx, M, dev, 1, c->st);
      launch_self_force(c->seg.as<double>(), c->resLd, dr, (int)nIdx, self_consts(*k), dev, c->st);
      launch_rows(c, n, dr, nIdx, *k, dev, fl);
      HIPCHK(hipGetLastError());
      check_flags(c);
      store6(c, dev, nIdx, &r, out, outCols, flags);
    }
    c->resN = -1;
  });
}

int ddd_sigma_tilde(ddd_ctx* c, const ddd_consts* k, const ddd_network* net, const int64_t* idx,
                    int64_t nIdx, const double* x0, int64_t P, double* sigma, int flags) {
  return guard(c, "ddd_sigma_tilde", [&] {
    if (!k) bad("consts is NULL");
    check_net(net);
    if (P < 0 || (P > 0 && (!x0 || !sigma))) bad("points/sigma inconsistent");
    if (P > ((int64_t)1 << 34)) bad("too many points");
    std::vector<int64_t> r;
    if (idx) r = rows_from(idx, nIdx, net->numSeg);
    const int64_t m = idx ? nIdx : net->numSeg;
    const int64_t ld = std::max<int64_t>(m, 1);
    // Pack (n1, n2, b) of the selected segments; tangents on the device.
    c->h.assign((size_t)(12 * ld), 0.0);
    double* h = c->h.data();
    const int64_t L = net->segIdxLd;
    for (int64_t q = 0; q < m; ++q) {
      const int64_t i = idx ? r[(size_t)q] : q;
      const int64_t lk = net->segIdx[i], a = net->segIdx[i + L], b = net->segIdx[i + 2 * L];
      if (lk < 1 || lk > net->bVecCols) bad("segIdx[:,1] outside bVec");
      if (a < 1 || a > net->coordCols || b < 1 || b > net->coordCols)
        bad("segIdx node outside coord");
      for (int e = 0; e < 3; ++e) {
        h[(0 + e) * ld + q] = net->coord[3 * (a - 1) + e];
        h[(3 + e) * ld + q] = net->coord[3 * (b - 1) + e];
        h[(6 + e) * ld + q] = net->bVec[3 * (lk - 1) + e];
      }
    }
    c->sseg.ensure(sizeof(double) * kSigmaFields * ld);
    HIPCHK(hipMemcpyAsync(c->sseg.p, h, sizeof(double) * 12 * ld, hipMemcpyHostToDevice, c->st));
    const SigmaK K{k->coreRadSq, k->mu4pi, k->mu4pinu, k->nuMu4pinu, k->coreRadSq * k->mu8pi};
    launch_prep_sigma_segments(c->sseg.as<double>(), ld, (int)m, K, c->st);
    if (P == 0) {
      HIPCHK(hipStreamSynchronize(c->st));
      return;
    }
    c->pts.ensure(sizeof(double) * 3 * P);
    HIPCHK(hipMemcpyAsync(c->pts.p, x0, sizeof(double) * 3 * P, hipMemcpyHostToDevice, c->st));
    const bool acc = flags & DDD_ACCUMULATE;
    c->sigma.ensure(sizeof(double) * 6 * P);
    if (acc)
      HIPCHK(hipMemcpyAsync(c->sigma.p, sigma, sizeof(double) * 6 * P, hipMemcpyHostToDevice, c->st));
    const int nChunk = m > 0 ? sigma_chunks_for(P, (int)m) : 1;
    c->sigPart.ensure(sizeof(double) * 6 * P * nChunk);
    if (m > 0) {
      launch_sigma_tilde(c->sseg.as<double>(), ld, (int)m, c->pts.as<double>(), P, K,
                         c->sigPart.as<double>(), nChunk, c->st);
    } else {
      HIPCHK(hipMemsetAsync(c->sigPart.p, 0, sizeof(double) * 6 * P, c->st));
    }
    launch_sigma_reduce(c->sigPart.as<double>(), nChunk, P, c->sigma.as<double>(), acc ? 1 : 0,
                        c->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(sigma, c->sigma.p, sizeof(double) * 6 * P, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

int ddd_sigma_tilde_noalloc(ddd_ctx* c, const ddd_consts* k, int64_t numSeg, const int64_t* segIdx,
                            int64_t segIdxLd, const double* coord, int64_t coordCols,
                            const double* bVec, int64_t bVecCols, const int64_t* idx, int64_t nIdx,
                            const double* x0, int64_t P, double* sigma, int flags) {
  return guard(c, "ddd_sigma_tilde_noalloc", [&] {
    if (!k) bad("consts is NULL");
    if (numSeg < 0 || numSeg > segIdxLd) bad("numSeg outside [0, segIdxLd]");
    if (numSeg > (int64_t)1 << 30) bad("numSeg too large");
    if (numSeg > 0 && (!segIdx || !coord || !bVec)) bad("segIdx, coord and bVec are required");
    if (coordCols < 0 || bVecCols < 0) bad("negative column counts");
    if (nIdx < 0 || (nIdx > 0 && !idx) || nIdx > (int64_t)1 << 30) bad("idx/nIdx inconsistent");
    if (P < 0 || (P > 0 && (!x0 || !sigma))) bad("points/sigma inconsistent");
    const int64_t m = idx ? nIdx : numSeg;
    const int64_t ld = std::max<int64_t>(m, 1);
    if (P > ((int64_t)1 << 34)) bad("too many points");
    // scratch: the segment SoA, then (256-byte aligned) the chunk partials;
    // nChunk as ddd_sigma_tilde picks it (same sums, same bits), halved
    // until it fits
    const int64_t cap = kNoAllocScratchDoubles;
    const int64_t soa = ((int64_t)kSigmaFields * ld + 31) / 32 * 32;
    if (soa > cap || (P > 0 && m > 0 && (cap - soa) / 6 < P))
      bad("segments and points exceed the 64 MiB scratch (33 x segments + 6 x points doubles "
          "must fit in 2^23)");
    double* base = noalloc_scratch(c->device);
    if (!base) throw Err{DDD_ERR_HIP, "the scratch array's device address is unavailable"};
    if (P == 0) return;
    int nChunk = m > 0 ? sigma_chunks_for(P, (int)m) : 0;
    while (nChunk > 1 && soa + 6 * P * (int64_t)nChunk > cap) nChunk /= 2;
    double* sseg = base;
    double* part = base + soa;
    const SigmaK K{k->coreRadSq, k->mu4pi, k->mu4pinu, k->nuMu4pinu, k->coreRadSq * k->mu8pi};
    if (m > 0) {
      launch_gather_sigma_segments(segIdx, segIdxLd, numSeg, coord, coordCols, bVec, bVecCols, idx,
                                   (int)m, sseg, ld, c->st);
      launch_prep_sigma_segments(sseg, ld, (int)m, K, c->st);
      launch_sigma_tilde(sseg, ld, (int)m, x0, P, K, part, nChunk, c->st);
    }
    // nChunk 0 (no segments): sigma = 0, or unchanged when accumulating
    launch_sigma_reduce(part, nChunk, P, sigma, (flags & DDD_ACCUMULATE) ? 1 : 0, c->st);
    HIPCHK(hipGetLastError());
  });
}

int ddd_node_force(ddd_ctx* c, const ddd_consts* k, const ddd_network* net, const double* segForce,
                   int64_t segForceCols, double* nodeForce) {
  return guard(c, "ddd_node_force", [&] {
    if (!k || !net || !segForce || !nodeForce) bad("NULL argument");
    if (!net->connectivity || !net->links || !net->coord) bad("connectivity, links, coord required");
    const int64_t nn = net->numNode;
    if (nn < 0 || nn > net->connCols || nn > net->coordCols) bad("numNode outside arrays");
    if (net->connLd < 1) bad("connectivity leading dimension");
    const int64_t maxC = (net->connLd - 1) / 2;
    for (int64_t q = 0; q < nn; ++q) {
      const int64_t* cn = net->connectivity + net->connLd * q;
      if (cn[0] < 0 || cn[0] > maxC) bad("connectivity count");
      for (int64_t e = 1; e <= cn[0]; ++e) {
        const int64_t lk = cn[2 * e - 1], col = cn[2 * e];
        if (lk < 1 || lk > net->linksCols || lk > segForceCols) bad("connectivity link");
        if (col != 1 && col != 2) bad("connectivity column");
        const int64_t other = net->links[2 * (lk - 1) + (2 - col)];
        if (other < 1 || other > net->coordCols) bad("links node outside coord");
      }
    }
    if (nn == 0) return;
    c->conn.ensure(sizeof(int64_t) * net->connLd * nn);
    c->links.ensure(sizeof(int64_t) * 2 * net->linksCols);
    c->coord.ensure(sizeof(double) * 3 * net->coordCols);
    c->out.ensure(sizeof(double) * 6 * segForceCols);
    c->nodeF.ensure(sizeof(double) * 3 * nn);
    HIPCHK(hipMemcpyAsync(c->conn.p, net->connectivity, sizeof(int64_t) * net->connLd * nn,
                          hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->links.p, net->links, sizeof(int64_t) * 2 * net->linksCols,
                          hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->coord.p, net->coord, sizeof(double) * 3 * net->coordCols,
                          hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->out.p, segForce, sizeof(double) * 6 * segForceCols,
                          hipMemcpyHostToDevice, c->st));
    launch_node_gather(c->conn.as<int64_t>(), net->connLd, nn, c->links.as<int64_t>(),
                       c->coord.as<double>(), c->out.as<double>(), k->sigmaPN,
                       c->nodeF.as<double>(), c->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(nodeForce, c->nodeF.p, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost,
                          c->st));
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

int ddd_mobility_bcc(ddd_ctx* c, const ddd_consts* k, const ddd_drag* drag, const ddd_network* net,
                     const double* segForce, int64_t segForceCols, const int64_t* idx,
                     int64_t nIdx, double* nodeForce, double* nodeVel, int64_t outCols,
                     int flags) {
  return guard(c, "ddd_mobility_bcc", [&] {
    if (!k || !drag || !net || !segForce || !nodeForce || !nodeVel) bad("NULL argument");
    if (!net->connectivity || !net->links || !net->coord || !net->bVec)
      bad("connectivity, links, coord and bVec required");
    const int64_t nn = net->numNode;
    if (nn < 0 || nn > net->connCols || nn > net->coordCols) bad("numNode outside arrays");
    if (net->connLd < 1) bad("connectivity leading dimension");
    if (idx && nIdx < 0) bad("nIdx < 0");
    const int64_t n = idx ? nIdx : nn;
    const bool byPos = (flags & DDD_OUT_BY_POSITION) != 0;
    const int64_t maxC = (net->connLd - 1) / 2;
    for (int64_t q = 0; q < n; ++q) {
      const int64_t node = idx ? idx[q] : q + 1;
      if (node < 1 || node > nn) bad("idx outside 1..numNode");
      if (!byPos && node > outCols) bad("idx beyond the output columns");
      const int64_t* cn = net->connectivity + net->connLd * (node - 1);
      if (cn[0] < 0 || cn[0] > maxC) bad("connectivity count");
      for (int64_t e = 1; e <= cn[0]; ++e) {
        const int64_t lk = cn[2 * e - 1], col = cn[2 * e];
        if (lk < 1 || lk > net->linksCols || lk > segForceCols || lk > net->bVecCols)
          bad("connectivity link");
        if (col != 1 && col != 2) bad("connectivity column");
        const int64_t other = net->links[2 * (lk - 1) + (2 - col)];
        if (other < 1 || other > net->coordCols) bad("links node outside coord");
      }
    }
    if (byPos && outCols < n) bad("outCols < number of nodes");
    if (n == 0) return;
    c->conn.ensure(sizeof(int64_t) * net->connLd * nn);
    c->links.ensure(sizeof(int64_t) * 2 * net->linksCols);
    c->coord.ensure(sizeof(double) * 3 * net->coordCols);
    c->bvec.ensure(sizeof(double) * 3 * net->bVecCols);
    c->out.ensure(sizeof(double) * 6 * segForceCols);
    c->nodeF.ensure(sizeof(double) * 3 * n);
    c->nodeV.ensure(sizeof(double) * 3 * n);
    HIPCHK(hipMemcpyAsync(c->conn.p, net->connectivity, sizeof(int64_t) * net->connLd * nn,
                          hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->links.p, net->links, sizeof(int64_t) * 2 * net->linksCols,
                          hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->coord.p, net->coord, sizeof(double) * 3 * net->coordCols,
                          hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->bvec.p, net->bVec, sizeof(double) * 3 * net->bVecCols,
                          hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->out.p, segForce, sizeof(double) * 6 * segForceCols,
                          hipMemcpyHostToDevice, c->st));
    if (idx) {
      c->nodeIdx.ensure(sizeof(int64_t) * n);
      HIPCHK(hipMemcpyAsync(c->nodeIdx.p, idx, sizeof(int64_t) * n, hipMemcpyHostToDevice, c->st));
    }
    const double dr[4] = {drag->edge, drag->screw, drag->climb, drag->line};
    launch_mobility_bcc(c->conn.as<int64_t>(), net->connLd, c->links.as<int64_t>(),
                        c->coord.as<double>(), c->bvec.as<double>(), c->out.as<double>(),
                        idx ? c->nodeIdx.as<int64_t>() : nullptr, n, dr, k->sigmaPN, nullptr,
                        c->nodeF.as<double>(), c->nodeV.as<double>(), c->st);
    HIPCHK(hipGetLastError());
    if (byPos || !idx) {
      HIPCHK(hipMemcpyAsync(nodeForce, c->nodeF.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost,
                            c->st));
      HIPCHK(hipMemcpyAsync(nodeVel, c->nodeV.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost,
                            c->st));
      HIPCHK(hipStreamSynchronize(c->st));
      return;
    }
    // in place with a node list: columns idx[k] only
    c->h.resize(6 * n);
    HIPCHK(hipMemcpyAsync(c->h.data(), c->nodeF.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost,
                          c->st));
    HIPCHK(hipMemcpyAsync(c->h.data() + 3 * n, c->nodeV.p, sizeof(double) * 3 * n,
                          hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    for (int64_t q = 0; q < n; ++q) {
      const int64_t col = idx[q] - 1;
      for (int e = 0; e < 3; ++e) {
        nodeForce[3 * col + e] = c->h[3 * q + e];
        nodeVel[3 * col + e] = c->h[3 * n + 3 * q + e];
      }
    }
  });
}

namespace {

// Host geometry of calc_uTilde's external segments (DislocationFEM.jl:437-528):
// the point where each external end was projected out of the cuboid.
struct Box {
  double nrm[6][3], mid[6][3], lo[3], hi[3], centre[3];

  // convex-hull membership of a point computed on a face plane, with a
  // tolerance of 1e-9 of the box extent per axis
  bool contains(const double x[3]) const {
    for (int d = 0; d < 3; ++d) {
      const double tol = 1e-9 * (hi[d] - lo[d]);
      if (!(x[d] >= lo[d] - tol && x[d] <= hi[d] + tol)) return false;
    }
    return true;
  }
  static double dot(const double a[3], const double b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  }
  // findIntersectVolume (DislocationTopologySpecial.jl:7-38) through
  // linePlaneIntersect (Misc.jl:223-235): nearest in-volume intersection of
  // the line p + s l with a face plane; squared distance, Inf if none.
  double nearest(const double l[3], const double p[3], double out[3]) const {
    double best = std::numeric_limits<double>::infinity();
    out[0] = out[1] = out[2] = 0.0;
    for (int f = 0; f < 6; ++f) {
      const double den = dot(l, nrm[f]);
      if (den == 0.0) continue;
      const double diff[3] = {mid[f][0] - p[0], mid[f][1] - p[1], mid[f][2] - p[2]};
      const double s = dot(diff, nrm[f]) / den;
      const double x[3] = {p[0] + s * l[0], p[1] + s * l[1], p[2] + s * l[2]};
      if (std::isinf(x[0]) || std::isinf(x[1]) || std::isinf(x[2]) || !contains(x)) continue;
      const double r[3] = {x[0] - p[0], x[1] - p[1], x[2] - p[2]};
      const double d2 = std::min((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2], best);
      if (d2 < best) {
        best = d2;
        out[0] = x[0];
        out[1] = x[1];
        out[2] = x[2];
      }
    }
    return best;
  }
  // Projection of an external end: the nearest intersection along the
  // normals of faces 1, 3 and 5; failing that, steps over the face planes
  // towards the centroid (the reference adds the face midpoint in the plane
  // distance and shares the running minimum dc between both ends).
  void project(const double p[3], double& dc, double out[3]) const {
    double best = std::numeric_limits<double>::infinity();
    for (int j = 0; j < 6; j += 2) {
      double x[3];
      const double d = nearest(nrm[j], p, x);
      if (d < best) {
        best = d;
        out[0] = x[0];
        out[1] = x[1];
        out[2] = x[2];
      }
    }
    if (!std::isinf(best)) return;
    double t[3] = {p[0], p[1], p[2]};
    for (int f = 0; f < 6; ++f) {
      const double tm[3] = {t[0] + mid[f][0], t[1] + mid[f][1], t[2] + mid[f][2]};
      const double pd = dot(tm, nrm[f]);
      const double step = std::sqrt(pd * pd);
      const double x[3] = {t[0] - nrm[f][0] * step, t[1] - nrm[f][1] * step, t[2] - nrm[f][2] * step};
      const double r[3] = {x[0] - centre[0], x[1] - centre[1], x[2] - centre[2]};
      const double d = std::sqrt(dot(r, r));
      if (d <= dc) {
        dc = d;
        t[0] = x[0];
        t[1] = x[1];
        t[2] = x[2];
      }
    }
    out[0] = t[0];
    out[1] = t[1];
    out[2] = t[2];
  }
};

// Triangles of calc_uTilde in segment order: 13 x ld SoA (A, B, C, b, kind).
int64_t u_tilde_triangles(const ddd_network* net, const int32_t* label, const double* vertices,
                          const double* faceNorm, const double* faceMidPt,
                          std::vector<double>& tri, int64_t& ld) {
  Box bx;
  double cs[3] = {0, 0, 0};
  for (int d = 0; d < 3; ++d) {
    bx.lo[d] = std::numeric_limits<double>::infinity();
    bx.hi[d] = -std::numeric_limits<double>::infinity();
  }
  for (int v = 0; v < 8; ++v)
    for (int d = 0; d < 3; ++d) {
      cs[d] += vertices[3 * v + d];
      bx.lo[d] = std::min(bx.lo[d], vertices[3 * v + d]);
      bx.hi[d] = std::max(bx.hi[d], vertices[3 * v + d]);
    }
  for (int d = 0; d < 3; ++d) bx.centre[d] = cs[d] / 8.0;
  for (int f = 0; f < 6; ++f)
    for (int d = 0; d < 3; ++d) {
      bx.nrm[f][d] = faceNorm[3 * f + d];
      bx.mid[f][d] = faceMidPt[3 * f + d];
    }
  struct T {
    double a[3], b[3], c[3], bv[3], kind;
  };
  std::vector<T> out;
  out.reserve((size_t)net->numSeg);
  std::vector<int64_t> surfCon((size_t)net->coordCols + 1, 0);
  const int32_t srfMob = 3, srfFix = 4, ext = 5;
  auto xyz = [&](int64_t node, double r[3]) {
    for (int d = 0; d < 3; ++d) r[d] = net->coord[3 * (node - 1) + d];
  };
  for (int64_t i = 0; i < net->numSeg; ++i) {
    const int64_t n1 = net->links[2 * i], n2 = net->links[2 * i + 1];
    const int32_t l1 = label[n1 - 1], l2 = label[n2 - 1];
    if ((l1 == srfMob || l1 == srfFix) && l2 == ext) { surfCon[n2] = n1; continue; }
    if ((l2 == srfMob || l2 == srfFix) && l1 == ext) { surfCon[n1] = n2; continue; }
    T t{};
    xyz(n1, t.a);
    xyz(n2, t.b);
    for (int d = 0; d < 3; ++d) t.bv[d] = net->bVec[3 * i + d];
    if (l1 != ext && l2 != ext) {
      for (int d = 0; d < 3; ++d) t.c[d] = bx.centre[d];
      t.kind = 1.0;
      out.push_back(t);
      continue;
    }
    double dc = std::numeric_limits<double>::infinity();
    double Ap[3], Bp[3];
    bx.project(t.a, dc, Ap);
    bx.project(t.b, dc, Bp);
    if (surfCon[n1] != 0) xyz(surfCon[n1], Ap);
    if (surfCon[n2] != 0) xyz(surfCon[n2], Bp);
    double Cp[3];
    for (int d = 0; d < 3; ++d) Cp[d] = (t.a[d] + Bp[d]) / 2.0;
    const double* corners[5][3] = {{Ap, Bp, bx.centre}, {Ap, t.a, Cp}, {t.a, t.b, Cp},
                                   {t.b, Bp, Cp}, {Bp, Ap, Cp}};
    for (auto& q : corners) {
      T s{};
      for (int d = 0; d < 3; ++d) {
        s.a[d] = q[0][d];
        s.b[d] = q[1][d];
        s.c[d] = q[2][d];
        s.bv[d] = t.bv[d];
      }
      s.kind = 0.0;
      out.push_back(s);
    }
  }
  const int64_t n = (int64_t)out.size();
  ld = std::max<int64_t>(n, 1);
  tri.assign((size_t)(13 * ld), 0.0);
  for (int64_t q = 0; q < n; ++q) {
    const T& t = out[(size_t)q];
    for (int d = 0; d < 3; ++d) {
      tri[(0 + d) * ld + q] = t.a[d];
      tri[(3 + d) * ld + q] = t.b[d];
      tri[(6 + d) * ld + q] = t.c[d];
      tri[(9 + d) * ld + q] = t.bv[d];
    }
    tri[12 * ld + q] = t.kind;
  }
  return n;
}

}  // namespace

int ddd_u_tilde(ddd_ctx* c, double omNuInv8pi, double om2NuOmNuInv8pi, const ddd_network* net,
                const int32_t* label, const double* vertices, const double* faceNorm,
                const double* faceMidPt, const double* points, int64_t P, double* u, int flags) {
  return guard(c, "ddd_u_tilde", [&] {
    if (!net || !vertices || !faceNorm || !faceMidPt) bad("NULL argument");
    if (!label && net->numSeg > 0) bad("label is NULL");
    if (P < 0) bad("P < 0");
    if (P > 0 && (!points || !u)) bad("points and u required");
    if (!net->links || !net->coord || !net->bVec) bad("links, coord and bVec required");
    if (net->numSeg < 0 || net->numSeg > net->linksCols || net->numSeg > net->bVecCols)
      bad("numSeg outside links or bVec");
    if (P > ((int64_t)1 << 31)) bad("P too large");
    for (int64_t i = 0; i < net->numSeg; ++i)
      for (int e = 0; e < 2; ++e) {
        const int64_t nd = net->links[2 * i + e];
        if (nd < 1 || nd > net->coordCols) bad("links node outside coord");
      }
    std::vector<double> tri;
    int64_t ld = 0;
    const int64_t ntri = u_tilde_triangles(net, label, vertices, faceNorm, faceMidPt, tri, ld);
    if (ntri > ((int64_t)1 << 30)) bad("too many triangles");
    if (P == 0) return;
    const bool acc = (flags & DDD_ACCUMULATE) != 0;
    c->uOut.ensure(sizeof(double) * 3 * P);
    if (acc) {
      HIPCHK(hipMemcpyAsync(c->uOut.p, u, sizeof(double) * 3 * P, hipMemcpyHostToDevice, c->st));
    }
    if (ntri == 0) {
      if (!acc) HIPCHK(hipMemsetAsync(c->uOut.p, 0, sizeof(double) * 3 * P, c->st));
    } else {
      const int nChunk = u_tilde_chunks_for(P, (int)ntri);
      c->uTri.ensure(sizeof(double) * tri.size());
      c->uPts.ensure(sizeof(double) * 3 * P);
      c->uPart.ensure(sizeof(double) * 3 * P * nChunk);
      HIPCHK(hipMemcpyAsync(c->uTri.p, tri.data(), sizeof(double) * tri.size(),
                            hipMemcpyHostToDevice, c->st));
      HIPCHK(hipMemcpyAsync(c->uPts.p, points, sizeof(double) * 3 * P, hipMemcpyHostToDevice,
                            c->st));
      launch_u_tilde(c->uTri.as<double>(), ld, (int)ntri, c->uPts.as<double>(), P, omNuInv8pi,
                     om2NuOmNuInv8pi, c->uPart.as<double>(), nChunk, c->st);
      launch_chunk_sum(c->uPart.as<double>(), nChunk, 3 * P, c->uOut.as<double>(), acc ? 1 : 0,
                       c->st);
      HIPCHK(hipGetLastError());
    }
    HIPCHK(hipMemcpyAsync(u, c->uOut.p, sizeof(double) * 3 * P, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));  // tri (host) is released on return
  });
}

int ddd_detect_collision(ddd_ctx* c, double collisionDistSq, const ddd_network* net,
                         const int32_t* label, const double* nodeVel, const int64_t* skipSegs,
                         int64_t nSkip, ddd_collision* out) {
  return guard(c, "ddd_detect_collision", [&] {
    if (!net || !out) bad("NULL argument");
    const int64_t nn = net->numNode, ns = net->numSeg;
    if (nn < 0 || ns < 0) bad("negative sizes");
    if ((nn > 0 || ns > 0) && (!label || !nodeVel || !net->coord || !net->links))
      bad("label, nodeVel, coord and links required");
    if (nn > 0 && !net->connectivity) bad("connectivity required");
    if (ns > 0 && !net->bVec) bad("bVec required");
    if (nSkip < 0 || (nSkip > 0 && !skipSegs)) bad("skipSegs/nSkip inconsistent");
    if (nn > net->connCols || nn > net->coordCols) bad("numNode outside arrays");
    if (ns > net->linksCols || ns > net->bVecCols) bad("numSeg outside links or bVec");
    if (ns > INT32_MAX) bad("numSeg too large");
    for (int64_t i = 0; i < ns; ++i)
      for (int e = 0; e < 2; ++e) {
        const int64_t nd = net->links[2 * i + e];
        if (nd < 1 || nd > net->coordCols) bad("links node outside coord");
      }
    const int64_t maxC = net->connLd >= 1 ? (net->connLd - 1) / 2 : 0;
    for (int64_t q = 0; q < nn; ++q) {
      const int64_t* cn = net->connectivity + net->connLd * q;
      if (cn[0] < 0 || cn[0] > maxC) bad("connectivity count");
      for (int64_t e = 1; e <= cn[0]; ++e) {
        const int64_t lk = cn[2 * e - 1], col = cn[2 * e];
        if (lk < 1 || lk > net->linksCols) bad("connectivity link");
        if (col != 1 && col != 2) bad("connectivity column");
        const int64_t other = net->links[2 * (lk - 1) + (2 - col)];
        if (other < 1 || other > net->coordCols) bad("links node outside coord");
      }
    }
    *out = ddd_collision{0, 0, 0, 0, 0, 0, 0, 0, 0.0, 0.0};
    const int64_t cols = net->coordCols;
    c->cCoord.ensure(sizeof(double) * 3 * std::max<int64_t>(cols, 1));
    c->cVel.ensure(sizeof(double) * 3 * std::max<int64_t>(cols, 1));
    c->cLabel.ensure(sizeof(int32_t) * std::max<int64_t>(cols, 1));
    c->cLinks.ensure(sizeof(int64_t) * 2 * std::max<int64_t>(net->linksCols, 1));
    c->cSkip.ensure(sizeof(int64_t) * 2 * std::max<int64_t>(nSkip, 1));
    c->cRes.ensure(sizeof(CollisionCand) + sizeof(CollisionLast));
    if (cols > 0) {
      HIPCHK(hipMemcpyAsync(c->cCoord.p, net->coord, sizeof(double) * 3 * cols,
                            hipMemcpyHostToDevice, c->st));
      HIPCHK(hipMemcpyAsync(c->cVel.p, nodeVel, sizeof(double) * 3 * cols, hipMemcpyHostToDevice,
                            c->st));
      HIPCHK(hipMemcpyAsync(c->cLabel.p, label, sizeof(int32_t) * cols, hipMemcpyHostToDevice,
                            c->st));
    }
    if (net->linksCols > 0)
      HIPCHK(hipMemcpyAsync(c->cLinks.p, net->links, sizeof(int64_t) * 2 * net->linksCols,
                            hipMemcpyHostToDevice, c->st));
    // skip list as sorted keys min * 2^32 + max (binary-searched per pair)
    std::vector<int64_t> skipKeys((size_t)nSkip);
    for (int64_t q = 0; q < nSkip; ++q) {
      const int64_t a = skipSegs[2 * q], b = skipSegs[2 * q + 1];
      if (a < 0 || b < 0 || a >= ((int64_t)1 << 31) || b >= ((int64_t)1 << 31))
        bad("skipSegs entry outside [0, 2^31)");
      skipKeys[(size_t)q] = a < b ? (a << 32) + b : (b << 32) + a;
    }
    std::sort(skipKeys.begin(), skipKeys.end());
    if (nSkip > 0)
      HIPCHK(hipMemcpyAsync(c->cSkip.p, skipKeys.data(), sizeof(int64_t) * nSkip,
                            hipMemcpyHostToDevice, c->st));
    CollisionCand* resB = reinterpret_cast<CollisionCand*>(c->cRes.p);
    CollisionLast* resL = reinterpret_cast<CollisionLast*>(static_cast<char*>(c->cRes.p) +
                                                           sizeof(CollisionCand));
    CollisionCand hb{0.0, -1, -1, {0, 0, 0, 0, 0, 0}};
    CollisionLast hl{-1, -1, 0.0, 0.0};
    // hinges
    if (nn > 0) {
      c->cConn.ensure(sizeof(int64_t) * net->connLd * nn);
      c->cBest.ensure(sizeof(CollisionCand) * nn);
      c->cLast.ensure(sizeof(CollisionLast) * nn);
      HIPCHK(hipMemcpyAsync(c->cConn.p, net->connectivity, sizeof(int64_t) * net->connLd * nn,
                            hipMemcpyHostToDevice, c->st));
      launch_collision_hinge(nn, c->cConn.as<int64_t>(), net->connLd, c->cLinks.as<int64_t>(),
                             c->cCoord.as<double>(), c->cVel.as<double>(),
                             c->cLabel.as<int32_t>(), c->cSkip.as<int64_t>(), nSkip,
                             collisionDistSq, c->cBest.as<CollisionCand>(),
                             c->cLast.as<CollisionLast>(), c->st);
      launch_collision_final(c->cBest.as<CollisionCand>(), nn, c->cLast.as<CollisionLast>(), nn,
                             nullptr, 0, resB, resL, c->st);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(&hb, resB, sizeof(hb), hipMemcpyDeviceToHost, c->st));
      HIPCHK(hipMemcpyAsync(&hl, resL, sizeof(hl), hipMemcpyDeviceToHost, c->st));
      HIPCHK(hipStreamSynchronize(c->st));
    }
    auto fill = [&](const CollisionCand& b, int type, const CollisionLast& l) {
      out->collision = 1;
      out->type = type;
      out->n1s1 = b.info[0];
      out->n2s1 = b.info[1];
      out->n1s2 = b.info[2];
      out->n2s2 = b.info[3];
      out->s1 = b.info[4];
      out->s2 = b.info[5];
      out->L1 = l.L1;
      out->L2 = l.L2;
    };
    if (hb.a >= 0) {
      fill(hb, 1, hl);
      return;
    }
    // pairs of links without a shared node
    CollisionCand pb{0.0, -1, -1, {0, 0, 0, 0, 0, 0}};
    CollisionLast pl{-1, -1, 0.0, 0.0};
    if (ns >= 2) {
      const int64_t ld = ((ns + 63) / 64) * 64;
      const int blocks = collision_pair_blocks(ns);
      const int lastBlocks = collision_last_blocks(ns);
      c->cB.ensure(sizeof(double) * 3 * ns);
      c->cSeg.ensure(sizeof(double) * 19 * ld);
      c->cNodes.ensure(sizeof(int64_t) * 2 * ld);
      c->cFlags.ensure(sizeof(int) * ld);
      c->cBest.ensure(sizeof(CollisionCand) * blocks);
      c->cLast.ensure(sizeof(CollisionLast) * lastBlocks);
      HIPCHK(hipMemcpyAsync(c->cB.p, net->bVec, sizeof(double) * 3 * ns, hipMemcpyHostToDevice,
                            c->st));
      // links with both ends mobile, as a "largest such link <= j" index
      // (the last-evaluated scan hops over the others)
      std::vector<int> prevMob((size_t)ns);
      for (int64_t q = 0, last = -1; q < ns; ++q) {
        const int32_t la = label[net->links[2 * q] - 1], lb = label[net->links[2 * q + 1] - 1];
        if ((la == 1 || la == 3) && (lb == 1 || lb == 3)) last = q;
        prevMob[(size_t)q] = (int)last;
      }
      c->cPrev.ensure(sizeof(int) * ns);
      HIPCHK(hipMemcpyAsync(c->cPrev.p, prevMob.data(), sizeof(int) * ns, hipMemcpyHostToDevice,
                            c->st));
      launch_collision_pack(ns, ld, c->cLinks.as<int64_t>(), c->cCoord.as<double>(),
                            c->cVel.as<double>(), c->cB.as<double>(), c->cLabel.as<int32_t>(),
                            c->cSkip.as<int64_t>(), nSkip, c->cSeg.as<double>(),
                            c->cNodes.as<int64_t>(), c->cFlags.as<int>(), c->st);
      // scan order: Morton order of the midpoints over the nodes' bounding box
      double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {0.0, 0.0, 0.0};
      for (int64_t q = 0; q < ns; ++q)
        for (int e = 0; e < 2; ++e) {
          const double* x = net->coord + 3 * (net->links[2 * q + e] - 1);
          for (int d = 0; d < 3; ++d) {
            if ((q == 0 && e == 0) || x[d] < lo[d]) lo[d] = x[d];
            if ((q == 0 && e == 0) || x[d] > hi[d]) hi[d] = x[d];
          }
        }
      const double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]});
      const double scale = std::isfinite(ext) && ext > 0.0 ? 1024.0 / ext : 1.0;
      size_t sortBytes = 0;
      HIPCHK(collision_sort_bytes((int)ns, &sortBytes));
      const int64_t nbS = (ns + 63) / 64;
      c->cKeys.ensure(sizeof(uint32_t) * ns);
      c->cKeysOut.ensure(sizeof(uint32_t) * ns);
      c->cIdx.ensure(sizeof(int) * ns);
      c->cPerm.ensure(sizeof(int) * ns);
      c->cSortWork.ensure(std::max<size_t>(sortBytes, 1));
      c->cBox.ensure(sizeof(double) * 6 * nbS);
      HIPCHK(launch_collision_order(c->cSeg.as<double>(), ld, (int)ns, lo, scale,
                                    c->cKeys.as<uint32_t>(), c->cKeysOut.as<uint32_t>(),
                                    c->cIdx.as<int>(), c->cPerm.as<int>(), c->cSortWork.p, sortBytes,
                                    c->cBox.as<double>(), c->st));
      launch_collision_pairs(c->cSeg.as<double>(), ld, ns, c->cNodes.as<int64_t>(),
                             c->cFlags.as<int>(), c->cSkip.as<int64_t>(), nSkip, collisionDistSq,
                             c->cPerm.as<int>(), c->cBox.as<double>(), blocks,
                             c->cBest.as<CollisionCand>(), c->st);
      launch_collision_last(c->cSeg.as<double>(), ld, ns, c->cNodes.as<int64_t>(),
                            c->cFlags.as<int>(), c->cPrev.as<int>(), c->cSkip.as<int64_t>(), nSkip,
                            lastBlocks, c->cLast.as<CollisionLast>(), c->st);
      launch_collision_final(c->cBest.as<CollisionCand>(), blocks, c->cLast.as<CollisionLast>(),
                             lastBlocks, c->cSeg.as<double>(), ld, resB, resL, c->st);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(&pb, resB, sizeof(pb), hipMemcpyDeviceToHost, c->st));
      HIPCHK(hipMemcpyAsync(&pl, resL, sizeof(pl), hipMemcpyDeviceToHost, c->st));
      HIPCHK(hipStreamSynchronize(c->st));
    }
    const CollisionLast& l = pl.a >= 0 ? pl : hl;
    if (pb.a >= 0) {
      fill(pb, 2, l);
      return;
    }
    if (l.a >= 0) {
      out->L1 = l.L1;
      out->L2 = l.L2;
    }
  });
}

int ddd_net_upload(ddd_ctx* c, const ddd_consts* k, const ddd_network* net, const ddd_mesh* mesh) {
  return guard(c, "ddd_net_upload", [&] {
    if (!k) bad("consts is NULL");
    check_net(net);
    upload_segments(c, net);
    c->resMesh = mesh != nullptr;
    if (mesh) c->resM = upload_mesh(c, mesh);
    c->resK = *k;
    c->resN = net->numSeg;
    c->force.ensure(sizeof(double) * 6 * std::max<int64_t>(c->resN, 1));
    reset_flags(c);  // the pair stage sizes its chunk buffers itself
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

namespace {
void net_stage(ddd_ctx* c, int stage, int withPK) {
  if (c->resN < 0) bad("no network uploaded (ddd_net_upload)");
  if (withPK && !c->resMesh) bad("no mesh uploaded");
  const int64_t n = c->resN;
  double* dev = c->force.as<double>();
  if (stage == 0) {
    dev = zeroed(c, c->force, 6 * n);
    if (withPK) launch_pk_force(c->seg.as<double>(), c->resLd, nullptr, (int)n, c->resM, dev, 1, c->st);
    launch_self_force(c->seg.as<double>(), c->resLd, nullptr, (int)n, self_consts(c->resK), dev, c->st);
  } else if (stage == 1) {
    // no pair stage below two segments, as segseg_all
    if (is_hybrid(&c->opts)) {
      if (n >= 2) hybrid_pairs(c, c->resK, n, dev, c->flags.as<unsigned>(), c->opts);
    } else {
      segseg_tiles(c, c->resK, n, c->flags.as<unsigned>(), &c->opts, dev);
    }
  } else if (stage == 2) {
    if (is_hybrid(&c->opts)) {
      if (n >= 2) hybrid_apply(c, n, dev, c->opts);
    } else {
      segseg_finish(c, n, dev);
    }
  } else if (stage == 3) {
    if (c->dynNodes < 0) bad("no dynamics state (ddd_net_upload_dynamics)");
    launch_mobility_bcc(c->dConn.as<int64_t>(), c->dynConnLd, c->dLinks.as<int64_t>(),
                        c->dCoord.as<double>(), c->dBvec.as<double>(), dev, nullptr, c->dynNodes,
                        c->dynDrag, c->resK.sigmaPN, c->dLabel.as<int32_t>(),
                        c->dNodeF.as<double>(), c->dVel.as<double>(), c->st);
  } else {
    bad("stage must be 0, 1, 2 or 3");
  }
  HIPCHK(hipGetLastError());
}
}  // namespace

int ddd_net_seg_force(ddd_ctx* c, int withPK) {
  return guard(c, "ddd_net_seg_force", [&] {
    for (int s = 0; s < 3; ++s) net_stage(c, s, withPK);
  });
}

int ddd_net_stage(ddd_ctx* c, int stage, int withPK) {
  return guard(c, "ddd_net_stage", [&] { net_stage(c, stage, withPK); });
}

int ddd_net_download(ddd_ctx* c, double* out, int64_t outCols) {
  return guard(c, "ddd_net_download", [&] {
    if (c->resN < 0) bad("no network uploaded (ddd_net_upload)");
    check_out(out, outCols, c->resN);
    check_flags(c);
    store6(c, c->force.as<double>(), c->resN, nullptr, out, outCols, 0);
  });
}

// Device memory the context's own buffers hold (far fields and networks
// excluded).
uint64_t ctx_device_bytes(ddd_ctx* c) {
  const DevBuf* all[] = {&c->seg, &c->chunkMem, &c->force, &c->rows, &c->out, &c->flags,
                         &c->meshC, &c->meshCoord, &c->meshConn, &c->uHat, &c->sseg, &c->pts,
                         &c->sigPart, &c->sigma, &c->conn, &c->links, &c->coord, &c->nodeF,
                         &c->bvec, &c->nodeV, &c->nodeIdx, &c->uTri, &c->uPts, &c->uPart, &c->uOut,
                         &c->cConn, &c->cLinks, &c->cCoord, &c->cVel, &c->cB, &c->cLabel,
                         &c->cSkip, &c->cSeg, &c->cNodes, &c->cFlags, &c->cBest, &c->cLast,
                         &c->cRes, &c->cKeys, &c->cKeysOut, &c->cIdx, &c->cPerm, &c->cSortWork,
                         &c->cBox, &c->cPrev, &c->counter, &c->mixStage, &c->specCnt, &c->specOff,
                         &c->specWork, &c->specAt, &c->specFixN, &c->redoN, &c->dCoord,
                         &c->dCoord0, &c->dVel, &c->dVel0, &c->dNodeF, &c->dLinks, &c->dConn,
                         &c->dBvec, &c->dLabel, &c->dSegNodes, &c->dRedPart, &c->dRed};
  uint64_t b = 0;
  for (const DevBuf* d : all) b += d->bytes;
  return b;
}

void fill_stats(ddd_ctx* c, int64_t n, const ddd_opts* o, ddd_stats* st) {
  if (!st) return;
  st->deviceBytes = ctx_device_bytes(c);
  st->chunks = is_hybrid(o) ? 0 : (uint64_t)c->chunks;
  st->pairs = n > 1 ? (uint64_t)n * (uint64_t)(n - 1) / 2 : 0;
  st->pairsFP64 = st->pairs;
  if (is_hybrid(o) && n > 1) {
    unsigned long long e[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(e, o->farfield->counters.p, sizeof(e), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    st->pairs = e[0];
    st->pairsFP64 = e[1];
    return;
  }
  if (o && o->mode == DDD_MODE_MIXED && n > 1) {
    unsigned long long e = 0;
    HIPCHK(hipMemcpyAsync(&e, c->counter.as<unsigned long long>() + 1, sizeof(e),
                          hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    st->pairsFP64 = e;
  }
  st->tileMs = 0.0;
  if (!is_hybrid(o) && c->tileEvUsed > 0) {
    HIPCHK(hipStreamSynchronize(c->st));
    for (int q = 0; q < c->tileEvUsed; ++q) {
      float ms = 0.0f;
      HIPCHK(hipEventElapsedTime(&ms, c->tileEv[2 * q], c->tileEv[2 * q + 1]));
      st->tileMs += ms;
    }
  }
}

int ddd_segseg_force_opts(ddd_ctx* c, const ddd_consts* k, const ddd_network* net,
                          const ddd_opts* opts, double* out, int64_t outCols, int flags,
                          ddd_stats* stats) {
  return guard(c, "ddd_segseg_force_opts", [&] {
    if (!k) bad("consts is NULL");
    check_net(net);
    check_opts(opts);
    const int64_t n = net->numSeg;
    check_out(out, outCols, n);
    upload_segments(c, net);
    unsigned* fl = reset_flags(c);
    double* dev = zeroed(c, c->force, 6 * n);
    segseg_all(c, *k, n, dev, fl, opts);
    check_flags(c);
    fill_stats(c, n, opts, stats);
    store6(c, dev, n, nullptr, out, outCols, flags);
    c->resN = -1;
  });
}

int ddd_net_set_opts(ddd_ctx* c, const ddd_opts* o) {
  return guard(c, "ddd_net_set_opts", [&] {
    check_opts(o);
    c->opts = o ? *o : ddd_opts{DDD_MODE_FP64, 0.0, 0.0, nullptr};
  });
}

int ddd_net_stats(ddd_ctx* c, ddd_stats* st) {
  return guard(c, "ddd_net_stats", [&] {
    if (c->resN < 0) bad("no network uploaded (ddd_net_upload)");
    if (!st) bad("stats is NULL");
    fill_stats(c, c->resN, &c->opts, st);
  });
}

int ddd_farfield_create(ddd_ctx* c, int G, const double lo[3], double h, int nearRadius,
                        ddd_farfield** out) {
  return guard(c, "ddd_farfield_create", [&] {
    if (!out || !lo) bad("NULL argument");
    if (G < 2 || G > 32) bad("G must be in [2, 32]");
    if (!(h > 0)) bad("h must be positive");
    if (nearRadius < 0 || nearRadius >= G) bad("nearRadius must be in [0, G)");
    ddd_farfield* ff = new ddd_farfield();
    ff->g.G = G;
    ff->g.h = h;
    for (int a = 0; a < 3; ++a) ff->g.lo[a] = lo[a];
    ff->rn = nearRadius;
    ff->NP = fft_size_for(G);
    ff->vol = (int64_t)ff->NP * ff->NP * ff->NP;
    ff->volH = (int64_t)ff->NP * ff->NP * (ff->NP / 2 + 1);
    std::vector<float2> tw(ff->NP);
    for (int k = 0; k < ff->NP; ++k) {
      const double a = -2.0 * 3.14159265358979323846 * k / ff->NP;
      tw[k] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    std::vector<int> mrank, rankCell;
    morton_ranks(G, mrank, rankCell);
    const int nc = G * G * G;
    try {
      ff->tw.ensure(sizeof(float2) * ff->NP);
      HIPCHK(hipMemcpyAsync(ff->tw.p, tw.data(), sizeof(float2) * ff->NP, hipMemcpyHostToDevice, c->st));
      ff->mrank.ensure(sizeof(int) * nc);
      ff->rankCell.ensure(sizeof(int) * nc);
      HIPCHK(hipMemcpyAsync(ff->mrank.p, mrank.data(), sizeof(int) * nc, hipMemcpyHostToDevice, c->st));
      HIPCHK(hipMemcpyAsync(ff->rankCell.p, rankCell.data(), sizeof(int) * nc, hipMemcpyHostToDevice,
                            c->st));
      HIPCHK(hipStreamSynchronize(c->st));
    } catch (...) {
      ff->release();
      delete ff;
      throw;
    }
    *out = ff;
  });
}

int ddd_farfield_destroy(ddd_ctx* c, ddd_farfield* ff) {
  if (!ff) return DDD_OK;
  return guard(c, "ddd_farfield_destroy", [&] {
    HIPCHK(hipStreamSynchronize(c->st));
    if (c->opts.farfield == ff) c->opts = ddd_opts{DDD_MODE_FP64, 0.0, 0.0, nullptr};
    ff->release();
    delete ff;
  });
}

int ddd_farfield_set_operator(ddd_ctx* c, ddd_farfield* ff, const float* W) {
  return guard(c, "ddd_farfield_set_operator", [&] {
    if (!ff || !W) bad("NULL argument");
    const int S = 2 * ff->g.G - 1;
    const int64_t nW = (int64_t)S * S * S * kChan;
    DevBuf dW, scratch;
    try {
      dW.ensure(sizeof(float) * nW);
      scratch.ensure(sizeof(float2) * kChanPairs * ff->vol);
      ff->What.ensure(sizeof(float2) * kChan * ff->volH);
      HIPCHK(hipMemcpyAsync(dW.p, W, sizeof(float) * nW, hipMemcpyHostToDevice, c->st));
      launch_operator_spectrum(dW.as<float>(), ff->g, ff->rn, ff->tw.as<float2>(),
                               scratch.as<float2>(), ff->What.as<float2>(), c->st);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(c->st));
    } catch (...) {
      dW.release();
      scratch.release();
      throw;
    }
    dW.release();
    scratch.release();
    ff->haveOperator = true;
  });
}

int ddd_farfield_debug(ddd_ctx* c, ddd_farfield* ff, int what, void* out, int64_t bytes) {
  return guard(c, "ddd_farfield_debug", [&] {
    if (!ff || !out) bad("NULL argument");
    const int64_t nc = (int64_t)ff->g.G * ff->g.G * ff->g.G;
    const DevBuf* b = nullptr;
    int64_t need = 0;
    if (what == 0) { b = &ff->M; need = (int64_t)sizeof(float) * kMoments * nc; }
    else if (what == 1) { b = &ff->L; need = (int64_t)sizeof(float) * kLocal * nc; }
    else if (what == 2) { b = &ff->cellOf; need = bytes; }
    else bad("what must be 0, 1 or 2");
    if (!b->p) bad("no hybrid evaluation yet");
    if (bytes < need || (size_t)need > b->bytes) bad("size mismatch");
    HIPCHK(hipMemcpyAsync(out, b->p, need, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

namespace {

const unsigned short* wb(ddd_mlp* m, int64_t off) { return m->Wb.as<unsigned short>() + off; }
const float* pp(ddd_mlp* m, int64_t off) { return m->P.as<float>() + off; }

// Y = MLP(features of off3) for ns samples (off3 already on the device).
// Each product's epilogue adds the layer's bias and writes the activation.
void mlp_forward(ddd_mlp* m, int ns, hipStream_t st) {
  float* wk = m->gemmWork.as<float>();
  auto layer = [&](int64_t offB, unsigned short* H) {
    MlpEpilogue ep;
    ep.bias = pp(m, offB);
    ep.act = 1;
    ep.H = H;
    return ep;
  };
  mlp_features(m->off3.as<int>(), ns, m->aH, m->X0.as<unsigned short>(), st);
  mlp_gemm(m->X0.as<unsigned short>(), kMlpIn, 1, wb(m, kOffW1), kHidden, 1, m->Z1.as<float>(),
           kHidden, ns, kHidden, kMlpIn, 0, st, wk, kGemmWork,
           layer(kOffB1, m->H1.as<unsigned short>()));
  mlp_gemm(m->H1.as<unsigned short>(), kHidden, 1, wb(m, kOffW2), kHidden, 1, m->Z2.as<float>(),
           kHidden, ns, kHidden, kHidden, 0, st, wk, kGemmWork,
           layer(kOffB2, m->H2.as<unsigned short>()));
  mlp_gemm(m->H2.as<unsigned short>(), kHidden, 1, wb(m, kOffW3), kHidden, 1, m->Z3.as<float>(),
           kHidden, ns, kHidden, kHidden, 0, st, wk, kGemmWork,
           layer(kOffB3, m->H3.as<unsigned short>()));
  mlp_gemm(m->H3.as<unsigned short>(), kHidden, 1, wb(m, kOffW4), kMlpOutPad, 1, m->Y.as<float>(),
           kMlpOutPad, ns, kMlpOutPad, kHidden, 0, st, wk, kGemmWork, layer(kOffB4, nullptr));
}

// Gradient of the loss into m->G from dY (bf16, ns x kMlpOutPad, in m->dYb).
void mlp_backward(ddd_mlp* m, int ns, hipStream_t st) {
  float* G = m->G.as<float>();
  float* wk = m->gemmWork.as<float>();
  float* part = m->colPart.as<float>();
  unsigned short* dYb = m->dYb.as<unsigned short>();
  unsigned short* dZ = m->dZb.as<unsigned short>();
  float* dH = m->dH.as<float>();
  mlp_gemm(m->H3.as<unsigned short>(), 1, kHidden, dYb, kMlpOutPad, 1, G + kOffW4, kMlpOutPad,
           kHidden, kMlpOutPad, ns, 0, st, wk, kGemmWork);
  mlp_colsum(dYb, ns, kMlpOutPad, part, G + kOffB4, st);
  mlp_gemm(dYb, kMlpOutPad, 1, wb(m, kOffW4), 1, kMlpOutPad, dH, kHidden, ns, kHidden, kMlpOutPad,
           0, st, wk, kGemmWork);
  mlp_act_grad_colsum(dH, m->Z3.as<float>(), ns, kHidden, dZ, part, G + kOffB3, st);
  mlp_gemm(m->H2.as<unsigned short>(), 1, kHidden, dZ, kHidden, 1, G + kOffW3, kHidden, kHidden,
           kHidden, ns, 0, st, wk, kGemmWork);
  mlp_gemm(dZ, kHidden, 1, wb(m, kOffW3), 1, kHidden, dH, kHidden, ns, kHidden, kHidden, 0, st, wk, kGemmWork);
  mlp_act_grad_colsum(dH, m->Z2.as<float>(), ns, kHidden, dZ, part, G + kOffB2, st);
  mlp_gemm(m->H1.as<unsigned short>(), 1, kHidden, dZ, kHidden, 1, G + kOffW2, kHidden, kHidden,
           kHidden, ns, 0, st, wk, kGemmWork);
  mlp_gemm(dZ, kHidden, 1, wb(m, kOffW2), 1, kHidden, dH, kHidden, ns, kHidden, kHidden, 0, st, wk, kGemmWork);
  mlp_act_grad_colsum(dH, m->Z1.as<float>(), ns, kHidden, dZ, part, G + kOffB1, st);
  mlp_gemm(m->X0.as<unsigned short>(), 1, kMlpIn, dZ, kHidden, 1, G + kOffW1, kHidden, kMlpIn,
           kHidden, ns, 0, st, wk, kGemmWork);
}

void mlp_refresh_bf16(ddd_mlp* m, hipStream_t st) {
  mlp_to_bf16(m->P.as<float>(), kMlpParams, m->Wb.as<unsigned short>(), st);
}

float env_scale(const ddd_consts& k, const ddd_farfield* ff) { return (float)(k.mu4pi / ff->g.h); }

void operator_from_mlp(ddd_ctx* c, ddd_farfield* ff, ddd_mlp* m, const ddd_consts& k) {
  const int G = ff->g.G, S = 2 * G - 1;
  m->aH = (float)(k.coreRad / ff->g.h);
  // The network is evaluated at every offset of the grid, or (cubic) at the
  // canonical offsets only, whose operators the symmetry expands.
  const int no = m->cubic ? canonical_index(G - 1, G - 1, G - 1) + 1 : S * S * S;
  std::vector<int> off((size_t)3 * no);
  if (m->cubic) {
    for (int a = 0; a < G; ++a)
      for (int b = 0; b <= a; ++b)
        for (int cc = 0; cc <= b; ++cc) {
          const int o = canonical_index(a, b, cc);
          off[3 * o] = a;
          off[3 * o + 1] = b;
          off[3 * o + 2] = cc;
        }
  } else {
    for (int z = 0; z < S; ++z)
      for (int y = 0; y < S; ++y)
        for (int x = 0; x < S; ++x) {
          const int o = x + S * (y + S * z);
          off[3 * o] = x - (G - 1);
          off[3 * o + 1] = y - (G - 1);
          off[3 * o + 2] = z - (G - 1);
        }
  }
  ScopedBuf dW, dWc, scratch, dOff;
  const int chunk = 4096;
  dW.ensure(sizeof(float) * (int64_t)S * S * S * kChan);
  if (m->cubic) dWc.ensure(sizeof(float) * (int64_t)no * kChan);
  float* Wout = m->cubic ? dWc.as<float>() : dW.as<float>();
  dOff.ensure(sizeof(int) * 3 * no);
  HIPCHK(hipMemcpyAsync(dOff.p, off.data(), sizeof(int) * 3 * no, hipMemcpyHostToDevice, c->st));
  m->ensure_batch(chunk);
  for (int o0 = 0; o0 < no; o0 += chunk) {
    const int cnt = std::min(chunk, no - o0);
    HIPCHK(hipMemcpyAsync(m->off3.p, dOff.as<int>() + 3 * o0, sizeof(int) * 3 * cnt,
                          hipMemcpyDeviceToDevice, c->st));
    mlp_forward(m, cnt, c->st);
    mlp_operator_from_y(m->Y.as<float>(), m->off3.as<int>(), cnt, env_scale(k, ff),
                        Wout + (int64_t)o0 * kChan, c->st);
  }
  if (m->cubic) operator_expand(dWc.as<float>(), G, dW.as<float>(), c->st);
  scratch.ensure(sizeof(float2) * kChanPairs * ff->vol);
  ff->What.ensure(sizeof(float2) * kChan * ff->volH);
  launch_operator_spectrum(dW.as<float>(), ff->g, ff->rn, ff->tw.as<float2>(), scratch.as<float2>(),
                           ff->What.as<float2>(), c->st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->st));
  ff->haveOperator = true;
}

}  // namespace

int64_t ddd_mlp_num_params(void) { return kMlpParams; }

int ddd_mlp_create(ddd_ctx* c, uint64_t seed, ddd_mlp** out) {
  return guard(c, "ddd_mlp_create", [&] {
    if (!out) bad("NULL argument");
    std::vector<float> p((size_t)kMlpParams, 0.f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    auto fill = [&](int64_t off, int fin, int fout, double gain) {
      const double sd = gain / std::sqrt((double)fin);
      for (int64_t e = 0; e < (int64_t)fin * fout; ++e) p[(size_t)(off + e)] = (float)(sd * nd(rng));
    };
    fill(kOffW1, kMlpIn, kHidden, 1.7);
    fill(kOffW2, kHidden, kHidden, 1.7);
    fill(kOffW3, kHidden, kHidden, 1.7);
    fill(kOffW4, kHidden, kMlpOutPad, 0.05);
    ddd_mlp* m = new ddd_mlp();
    try {
      for (DevBuf* b : {&m->P, &m->G, &m->Mm, &m->Vv}) b->ensure(sizeof(float) * kMlpParams);
      m->Wb.ensure(sizeof(unsigned short) * kMlpParams);
      HIPCHK(hipMemcpyAsync(m->P.p, p.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice, c->st));
      HIPCHK(hipMemsetAsync(m->Mm.p, 0, sizeof(float) * kMlpParams, c->st));
      HIPCHK(hipMemsetAsync(m->Vv.p, 0, sizeof(float) * kMlpParams, c->st));
      mlp_refresh_bf16(m, c->st);
      HIPCHK(hipStreamSynchronize(c->st));
    } catch (...) {
      m->release();
      delete m;
      throw;
    }
    *out = m;
  });
}

int ddd_mlp_destroy(ddd_ctx* c, ddd_mlp* m) {
  if (!m) return DDD_OK;
  return guard(c, "ddd_mlp_destroy", [&] {
    HIPCHK(hipStreamSynchronize(c->st));
    m->release();
    delete m;
  });
}

int ddd_mlp_params(ddd_ctx* c, ddd_mlp* m, float* get, const float* set) {
  return guard(c, "ddd_mlp_params", [&] {
    if (!m) bad("NULL mlp");
    if (get)
      HIPCHK(hipMemcpyAsync(get, m->P.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, c->st));
    if (set) {
      HIPCHK(hipMemcpyAsync(m->P.p, set, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, c->st));
      mlp_refresh_bf16(m, c->st);
    }
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

int ddd_mlp_loss_grad(ddd_ctx* c, ddd_mlp* m, int ns, const int* off3, const float* M,
                      const float* Lt, const float* den, float envScale, double* loss,
                      float* grad) {
  return guard(c, "ddd_mlp_loss_grad", [&] {
    if (!m || !off3 || !M || !Lt || !den || !loss || !grad) bad("NULL argument");
    if (ns < 1 || ns > (1 << 20)) bad("ns out of range");
    for (int s = 0; s < ns; ++s)
      if (off3[3 * s] == 0 && off3[3 * s + 1] == 0 && off3[3 * s + 2] == 0) bad("zero offset");
    m->ensure_batch(ns);
    ScopedBuf dM, dL, dD;
    dM.ensure(sizeof(float) * kMoments * ns);
    dL.ensure(sizeof(float) * kLocal * ns);
    dD.ensure(sizeof(float) * ns);
    HIPCHK(hipMemcpyAsync(m->off3.p, off3, sizeof(int) * 3 * ns, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(dM.p, M, sizeof(float) * kMoments * ns, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(dL.p, Lt, sizeof(float) * kLocal * ns, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(dD.p, den, sizeof(float) * ns, hipMemcpyHostToDevice, c->st));
    mlp_forward(m, ns, c->st);
    mlp_contract(m->Y.as<float>(), m->off3.as<int>(), dM.as<float>(), dL.as<float>(), dD.as<float>(),
                 ns, envScale, m->dYb.as<unsigned short>(), m->lossPer.as<float>(), c->st);
    mlp_backward(m, ns, c->st);
    std::vector<float> lp((size_t)ns);
    HIPCHK(hipMemcpyAsync(lp.data(), m->lossPer.p, sizeof(float) * ns, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(grad, m->G.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    double L = 0;
    for (float v : lp) L += v;
    *loss = L;
  });
}

int ddd_farfield_set_operator_mlp(ddd_ctx* c, ddd_farfield* ff, ddd_mlp* m, const ddd_consts* k) {
  return guard(c, "ddd_farfield_set_operator_mlp", [&] {
    if (!ff || !m || !k) bad("NULL argument");
    operator_from_mlp(c, ff, m, *k);
  });
}

namespace {

// Pack net into seg, bin it on ff's grid and form its cell moments (in ff's
// buffers); returns the leading dimension.
int64_t bin_network(ddd_ctx* c, ddd_farfield* ff, const ddd_network* net, DevBuf& seg) {
  const int nc = ff->g.G * ff->g.G * ff->g.G;
  const int64_t n = net->numSeg;
  const int64_t ld = pack_segments(c, net, seg);
  ff->cellOf.ensure(sizeof(int) * std::max<int64_t>(n, 1));
  ff->tmp.ensure(sizeof(int) * std::max<int64_t>(n, 1));
  ff->order.ensure(sizeof(int) * std::max<int64_t>(n, 1));
  ff->counts.ensure(sizeof(unsigned) * nc);
  ff->cursor.ensure(sizeof(unsigned) * nc);
  ff->start.ensure(sizeof(int) * (nc + 1));
  ff->M.ensure(sizeof(float) * kMoments * nc);
  launch_cells(seg.as<double>(), ld, (int)n, ff->g, ff->cellOf.as<int>(), ff->counts.as<unsigned>(),
               ff->cursor.as<unsigned>(), ff->start.as<int>(), ff->tmp.as<int>(), ff->order.as<int>(),
               c->st);
  launch_moments(seg.as<double>(), ld, ff->g, ff->start.as<int>(), ff->order.as<int>(),
                 ff->M.as<float>(), c->st);
  HIPCHK(hipGetLastError());
  return ld;
}

// The 27-point tensor Gauss rule on the unit cell and the L2 projection onto
// the quadratic basis it defines (exact for the basis products).
Projection cell_projection() {
  Projection pj{};
  const double r = std::sqrt(0.6) / 2;
  const double xi[3] = {-r, 0.0, r}, w[3] = {5.0 / 18, 8.0 / 18, 5.0 / 18};
  for (int a = 0; a < 3; ++a) pj.xi[a] = xi[a];
  double B[27][kPoly], W[27];
  for (int p = 0; p < 27; ++p) {
    const double x = xi[p % 3], y = xi[(p / 3) % 3], z = xi[p / 9];
    const double v[kPoly] = {1, x, y, z, x * x, y * y, z * z, x * y, x * z, y * z};
    for (int l = 0; l < kPoly; ++l) B[p][l] = v[l];
    W[p] = w[p % 3] * w[(p / 3) % 3] * w[p / 9];
  }
  // Gram matrix and its inverse (Gauss-Jordan, symmetric positive definite).
  double Gm[kPoly][2 * kPoly] = {};
  for (int i = 0; i < kPoly; ++i) {
    for (int j = 0; j < kPoly; ++j)
      for (int p = 0; p < 27; ++p) Gm[i][j] += W[p] * B[p][i] * B[p][j];
    Gm[i][kPoly + i] = 1.0;
  }
  for (int col = 0; col < kPoly; ++col) {
    int piv = col;
    for (int i = col + 1; i < kPoly; ++i)
      if (std::fabs(Gm[i][col]) > std::fabs(Gm[piv][col])) piv = i;
    for (int j = 0; j < 2 * kPoly; ++j) std::swap(Gm[col][j], Gm[piv][j]);
    const double d = Gm[col][col];
    for (int j = 0; j < 2 * kPoly; ++j) Gm[col][j] /= d;
    for (int i = 0; i < kPoly; ++i) {
      if (i == col) continue;
      const double f = Gm[i][col];
      for (int j = 0; j < 2 * kPoly; ++j) Gm[i][j] -= f * Gm[col][j];
    }
  }
  for (int l = 0; l < kPoly; ++l)
    for (int p = 0; p < 27; ++p) {
      double a = 0;
      for (int j = 0; j < kPoly; ++j) a += Gm[l][kPoly + j] * B[p][j] * W[p];
      pj.P[l][p] = a;
    }
  return pj;
}

}  // namespace

int ddd_farfield_train(ddd_ctx* c, ddd_farfield* ff, ddd_mlp* m, const ddd_consts* k,
                       const ddd_network* net, const ddd_train_opts* o, ddd_train_report* rep) {
  return guard(c, "ddd_farfield_train", [&] {
    if (!ff || !m || !k || !o) bad("NULL argument");
    check_net(net);
    if (o->samples < 1 || o->steps < 1 || o->batch < 1 || o->batch > (1 << 16) || o->valSamples < 0)
      bad("bad training options");
    const int64_t NS = o->samples + o->valSamples;
    if (NS > INT32_MAX / kMoments) bad("too many samples");
    const auto t0 = std::chrono::steady_clock::now();
    const int G = ff->g.G, rn = ff->rn;
    if (rn + 1 > G - 1) bad("no far offsets on this grid");
    m->aH = (float)(k->coreRad / ff->g.h);
    m->cubic = o->plainOffsets ? 0 : 1;
    const int nc = G * G * G;
    // Training network: its own segment buffer (an uploaded network stays);
    // cells and moments in the farfield's buffers, which the next hybrid
    // evaluation recomputes.
    ScopedBuf tseg;
    const int64_t tld = bin_network(c, ff, net, tseg);
    std::vector<int> start(nc + 1);
    std::vector<float> Mh((size_t)kMoments * nc);
    HIPCHK(hipMemcpyAsync(start.data(), ff->start.p, sizeof(int) * (nc + 1), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(Mh.data(), ff->M.p, sizeof(float) * kMoments * nc, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    std::vector<int> live;
    for (int cc = 0; cc < nc; ++cc)
      if (start[cc + 1] > start[cc]) live.push_back(cc);
    if (live.empty()) bad("training network has no binned segments");
    // Samples: an occupied source cell and an offset outside the near
    // stencil, a share nearFrac of them within nearShells shells of it.
    const double nearFrac = o->nearFrac > 0 ? o->nearFrac : 0.5;
    const int nearShells = o->nearShells > 0 ? o->nearShells : 3;
    const double denFloor = o->denFloor > 0 ? o->denFloor : 1e-2;
    std::vector<int> off((size_t)3 * NS), src((size_t)NS);
    std::mt19937_64 rng(o->seed);
    std::uniform_int_distribution<size_t> pickLive(0, live.size() - 1);
    std::uniform_real_distribution<double> u01(0.0, 1.0);
    const int Rnear = std::min(G - 1, rn + nearShells);
    for (int64_t s = 0; s < NS; ++s) {
      const int R = (u01(rng) < nearFrac) ? Rnear : G - 1;
      std::uniform_int_distribution<int> pick(-R, R);
      int ox, oy, oz;
      do {
        ox = pick(rng);
        oy = pick(rng);
        oz = pick(rng);
      } while (std::max(std::abs(ox), std::max(std::abs(oy), std::abs(oz))) <= rn);
      off[3 * s] = ox;
      off[3 * s + 1] = oy;
      off[3 * s + 2] = oz;
      src[s] = live[pickLive(rng)];
    }
    // Exact local expansions (FP64 stresses, projected), on the device.
    ScopedBuf dSrc, dOff, dLt, dDen;
    dSrc.ensure(sizeof(int) * NS);
    dOff.ensure(sizeof(int) * 3 * NS);
    dLt.ensure(sizeof(float) * kLocal * NS);
    dDen.ensure(sizeof(float) * NS);
    HIPCHK(hipMemcpyAsync(dSrc.p, src.data(), sizeof(int) * NS, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(dOff.p, off.data(), sizeof(int) * 3 * NS, hipMemcpyHostToDevice, c->st));
    const SigK<double> SK{k->coreRadSq, k->mu4pi, k->mu4pinu, k->nuMu4pinu, k->coreRadSq * k->mu8pi};
    local_targets(tseg.as<double>(), tld, ff->g, ff->start.as<int>(), ff->order.as<int>(),
                  dSrc.as<int>(), dOff.as<int>(), (int)NS, SK, cell_projection(), dLt.as<float>(), c->st);
    HIPCHK(hipGetLastError());
    std::vector<float> Lt((size_t)kLocal * NS), den((size_t)NS);
    HIPCHK(hipMemcpyAsync(Lt.data(), dLt.p, sizeof(float) * kLocal * NS, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    // den = |L|^2 + (denFloor * typical |L| at that distance)^2, the typical
    // size being envScale |o|^-2 |M0| (order-0 moments of the source).
    const float es = env_scale(*k, ff);
    for (int64_t s = 0; s < NS; ++s) {
      double a = 0;
      for (int q = 0; q < kLocal; ++q) a += (double)Lt[kLocal * s + q] * Lt[kLocal * s + q];
      const double R2 = (double)off[3 * s] * off[3 * s] + (double)off[3 * s + 1] * off[3 * s + 1] +
                        (double)off[3 * s + 2] * off[3 * s + 2];
      double m0 = 0;
      for (int f = 0; f < 9; ++f) m0 += (double)Mh[(size_t)f * nc + src[s]] * Mh[(size_t)f * nc + src[s]];
      const double fl = denFloor * es / R2 * std::sqrt(m0);
      den[s] = (float)(a + fl * fl + 1e-30);
    }
    HIPCHK(hipMemcpyAsync(dDen.p, den.data(), sizeof(float) * NS, hipMemcpyHostToDevice, c->st));
    // Training: consecutive batches of a shuffled order of the training set,
    // gathered on the device; a new order each epoch (the tail is dropped).
    const int B = (int)std::min<int64_t>(o->batch, o->samples);
    m->ensure_batch(std::max<int>(B, 1024));
    ScopedBuf bM, bL, bD, dPerm, vIdx;
    bM.ensure(sizeof(float) * kMoments * B);
    bL.ensure(sizeof(float) * kLocal * B);
    bD.ensure(sizeof(float) * B);
    dPerm.ensure(sizeof(int) * o->samples);
    std::vector<int> perm((size_t)o->samples);
    for (int64_t s = 0; s < o->samples; ++s) perm[(size_t)s] = (int)s;
    std::vector<float> lp((size_t)std::max<int64_t>(B, 1));
    std::vector<double> losses;
    int64_t cursor = o->samples;
    const double b1 = 0.9, b2 = 0.999;
    for (int stp = 0; stp < o->steps; ++stp) {
      if (cursor + B > o->samples) {
        HIPCHK(hipStreamSynchronize(c->st));  // the previous order is no longer read
        std::shuffle(perm.begin(), perm.end(), rng);
        HIPCHK(hipMemcpyAsync(dPerm.p, perm.data(), sizeof(int) * o->samples, hipMemcpyHostToDevice,
                              c->st));
        cursor = 0;
      }
      gather_batch(dPerm.as<int>() + cursor, B, dOff.as<int>(), dSrc.as<int>(), ff->M.as<float>(), nc,
                   dLt.as<float>(), dDen.as<float>(), m->cubic, m->off3.as<int>(), bM.as<float>(),
                   bL.as<float>(), bD.as<float>(), c->st);
      cursor += B;
      mlp_forward(m, B, c->st);
      mlp_contract(m->Y.as<float>(), m->off3.as<int>(), bM.as<float>(), bL.as<float>(), bD.as<float>(),
                   B, es, m->dYb.as<unsigned short>(), m->lossPer.as<float>(), c->st);
      mlp_backward(m, B, c->st);
      ++m->step;
      const double frac = (double)stp / std::max(1, o->steps - 1);
      const double lr = o->lr * std::pow(o->lrFinal / o->lr, frac);
      mlp_adam(m->P.as<float>(), m->G.as<float>(), m->Mm.as<float>(), m->Vv.as<float>(), kMlpParams,
               (float)lr, (float)b1, (float)b2, (float)(1 - std::pow(b1, (double)m->step)),
               (float)(1 - std::pow(b2, (double)m->step)), 1e-8f, m->Wb.as<unsigned short>(),
               c->st);
      if (stp < 50 || stp >= o->steps - 50) {
        HIPCHK(hipMemcpyAsync(lp.data(), m->lossPer.p, sizeof(float) * B, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        double Ls = 0;
        for (int q = 0; q < B; ++q) Ls += lp[(size_t)q];
        losses.push_back(Ls);
      }
      HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(c->st));
    // Held-out error of the predicted local expansions, overall and per shell.
    double num = 0, dnm = 0, numN = 0, dnmN = 0, numS[8] = {0}, dnmS[8] = {0};
    std::vector<int> vi((size_t)B);
    vIdx.ensure(sizeof(int) * B);
    for (int64_t v0 = o->samples; v0 < NS; v0 += B) {
      const int cnt = (int)std::min<int64_t>(B, NS - v0);
      for (int q = 0; q < cnt; ++q) vi[(size_t)q] = (int)(v0 + q);
      HIPCHK(hipMemcpyAsync(vIdx.p, vi.data(), sizeof(int) * cnt, hipMemcpyHostToDevice, c->st));
      gather_batch(vIdx.as<int>(), cnt, dOff.as<int>(), dSrc.as<int>(), ff->M.as<float>(), nc,
                   dLt.as<float>(), dDen.as<float>(), m->cubic, m->off3.as<int>(), bM.as<float>(),
                   bL.as<float>(), bD.as<float>(), c->st);
      mlp_forward(m, cnt, c->st);
      mlp_contract(m->Y.as<float>(), m->off3.as<int>(), bM.as<float>(), bL.as<float>(), bD.as<float>(),
                   cnt, es, m->dYb.as<unsigned short>(), m->lossPer.as<float>(), c->st);
      HIPCHK(hipMemcpyAsync(lp.data(), m->lossPer.p, sizeof(float) * cnt, hipMemcpyDeviceToHost, c->st));
      HIPCHK(hipStreamSynchronize(c->st));
      for (int q = 0; q < cnt; ++q) {
        const int64_t s = v0 + q;
        double a = 0;
        for (int w = 0; w < kLocal; ++w) a += (double)Lt[kLocal * s + w] * Lt[kLocal * s + w];
        const double e2 = (double)lp[(size_t)q] * den[s] * cnt;  // |L_hat - L|^2
        num += e2;
        dnm += a;
        const int ch = std::max(std::abs(off[3 * s]), std::max(std::abs(off[3 * s + 1]), std::abs(off[3 * s + 2])));
        if (ch <= rn + 2) { numN += e2; dnmN += a; }
        const int sh = std::min(7, ch - rn - 1);
        numS[sh] += e2;
        dnmS[sh] += a;
      }
    }
    operator_from_mlp(c, ff, m, *k);
    if (rep) {
      const size_t h = std::min<size_t>(50, losses.size());
      double a = 0, b = 0;
      for (size_t q = 0; q < h; ++q) a += losses[q];
      for (size_t q = losses.size() - h; q < losses.size(); ++q) b += losses[q];
      rep->lossFirst = h ? a / h : 0;
      rep->lossLast = h ? b / h : 0;
      rep->relErrVal = dnm > 0 ? std::sqrt(num / dnm) : 0;
      rep->relErrValNear = dnmN > 0 ? std::sqrt(numN / dnmN) : 0;
      rep->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      for (int q = 0; q < 8; ++q) rep->relErrShell[q] = dnmS[q] > 0 ? std::sqrt(numS[q] / dnmS[q]) : 0;
    }
  });
}

int ddd_farfield_local_targets(ddd_ctx* c, ddd_farfield* ff, const ddd_consts* k,
                               const ddd_network* net, int ns, const int* srcCell,
                               const int* off3, float* out) {
  return guard(c, "ddd_farfield_local_targets", [&] {
    if (!ff || !k || !srcCell || !off3 || !out) bad("NULL argument");
    check_net(net);
    if (ns < 1 || ns > (1 << 24)) bad("ns out of range");
    const int nc = ff->g.G * ff->g.G * ff->g.G;
    for (int s = 0; s < ns; ++s)
      if (srcCell[s] < 0 || srcCell[s] >= nc) bad("source cell outside the grid");
    ScopedBuf tseg, dSrc, dOff, dLt;
    const int64_t tld = bin_network(c, ff, net, tseg);
    dSrc.ensure(sizeof(int) * ns);
    dOff.ensure(sizeof(int) * 3 * ns);
    dLt.ensure(sizeof(float) * kLocal * ns);
    HIPCHK(hipMemcpyAsync(dSrc.p, srcCell, sizeof(int) * ns, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(dOff.p, off3, sizeof(int) * 3 * ns, hipMemcpyHostToDevice, c->st));
    const SigK<double> SK{k->coreRadSq, k->mu4pi, k->mu4pinu, k->nuMu4pinu, k->coreRadSq * k->mu8pi};
    local_targets(tseg.as<double>(), tld, ff->g, ff->start.as<int>(), ff->order.as<int>(),
                  dSrc.as<int>(), dOff.as<int>(), ns, SK, cell_projection(), dLt.as<float>(), c->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, dLt.p, sizeof(float) * kLocal * ns, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

const double* ddd_net_segforce_device(ddd_ctx* c) {
  return (c && c->resN >= 0) ? c->force.as<double>() : nullptr;
}

namespace {

// Julia's min/max for Float64: NaN wins, and -0.0 < 0.0.
double jmin(double a, double b) {
  if (std::isnan(a) || std::isnan(b)) return a + b;
  return (b < a || (std::signbit(b) && !std::signbit(a))) ? b : a;
}
double jmax(double a, double b) {
  if (std::isnan(a) || std::isnan(b)) return a + b;
  return (b > a || (!std::signbit(b) && std::signbit(a))) ? b : a;
}

void dyn_deriv(ddd_ctx* c, int withPK) {
  if (c->dynNodes < 0) bad("no dynamics state (ddd_net_upload_dynamics)");
  const int64_t ns = c->resN;
  launch_gather_coords(c->seg.as<double>(), c->resLd, ns, c->dSegNodes.as<int64_t>(),
                       c->dCoord.as<double>(), c->st);
  launch_prep_segments(c->seg.as<double>(), c->resLd, ns, c->st);
  for (int s = 0; s < 4; ++s) net_stage(c, s, withPK);
}

}  // namespace

int ddd_net_upload_dynamics(ddd_ctx* c, const ddd_network* net, const int32_t* label,
                            const ddd_drag* drag) {
  return guard(c, "ddd_net_upload_dynamics", [&] {
    if (!net || !label || !drag) bad("NULL argument");
    if (c->resN < 0 || c->resN != net->numSeg)
      bad("upload the same network with ddd_net_upload first");
    if (!net->connectivity || !net->links || !net->coord || !net->bVec || !net->segIdx)
      bad("segIdx, coord, bVec, links and connectivity required");
    const int64_t nn = net->numNode, ns = net->numSeg, L = net->segIdxLd;
    if (nn < 0 || nn > net->connCols || nn > net->coordCols) bad("numNode outside arrays");
    if (ns > net->linksCols || ns > net->bVecCols) bad("numSeg outside links or bVec");
    if (net->connLd < 1) bad("connectivity leading dimension");
    std::vector<int64_t> nodes((size_t)std::max<int64_t>(2 * ns, 1));
    for (int64_t i = 0; i < ns; ++i) {
      if (net->segIdx[i] != i + 1) bad("segIdx[:, 1] must be 1:numSeg, as getSegmentIdx! makes it");
      const int64_t a = net->segIdx[i + L], b = net->segIdx[i + 2 * L];
      if (a < 1 || a > nn || b < 1 || b > nn) bad("segment node outside 1:numNode");
      nodes[2 * i] = a - 1;
      nodes[2 * i + 1] = b - 1;
    }
    const int64_t maxC = (net->connLd - 1) / 2;
    for (int64_t q = 0; q < nn; ++q) {
      if (label[q] == 3)
        bad("srfMobDln nodes: the reference's surface velocity correction cannot run");
      const int64_t* cn = net->connectivity + net->connLd * q;
      if (cn[0] < 0 || cn[0] > maxC) bad("connectivity count");
      for (int64_t e = 1; e <= cn[0]; ++e) {
        const int64_t lk = cn[2 * e - 1], col = cn[2 * e];
        if (lk < 1 || lk > ns) bad("connectivity link outside 1:numSeg");
        if (col != 1 && col != 2) bad("connectivity column");
        const int64_t other = net->links[2 * (lk - 1) + (2 - col)];
        if (other < 1 || other > nn) bad("links node outside 1:numNode");
      }
    }
    const size_t m = (size_t)std::max<int64_t>(3 * nn, 1);
    c->dCoord.ensure(sizeof(double) * m);
    c->dCoord0.ensure(sizeof(double) * m);
    c->dVel.ensure(sizeof(double) * m);
    c->dVel0.ensure(sizeof(double) * m);
    c->dNodeF.ensure(sizeof(double) * m);
    c->dLinks.ensure(sizeof(int64_t) * std::max<int64_t>(2 * ns, 1));
    c->dConn.ensure(sizeof(int64_t) * std::max<int64_t>(net->connLd * nn, 1));
    c->dBvec.ensure(sizeof(double) * std::max<int64_t>(3 * ns, 1));
    c->dLabel.ensure(sizeof(int32_t) * std::max<int64_t>(nn, 1));
    c->dSegNodes.ensure(sizeof(int64_t) * nodes.size());
    c->dRedPart.ensure(sizeof(unsigned long long) * 2 * trapezoid_parts());
    c->dRed.ensure(sizeof(double) * 2);
    if (nn > 0) {
      HIPCHK(hipMemcpyAsync(c->dCoord.p, net->coord, sizeof(double) * 3 * nn,
                            hipMemcpyHostToDevice, c->st));
      HIPCHK(hipMemcpyAsync(c->dConn.p, net->connectivity, sizeof(int64_t) * net->connLd * nn,
                            hipMemcpyHostToDevice, c->st));
      HIPCHK(hipMemcpyAsync(c->dLabel.p, label, sizeof(int32_t) * nn, hipMemcpyHostToDevice,
                            c->st));
      HIPCHK(hipMemsetAsync(c->dVel.p, 0, sizeof(double) * 3 * nn, c->st));
      HIPCHK(hipMemsetAsync(c->dNodeF.p, 0, sizeof(double) * 3 * nn, c->st));
    }
    if (ns > 0) {
      HIPCHK(hipMemcpyAsync(c->dLinks.p, net->links, sizeof(int64_t) * 2 * ns,
                            hipMemcpyHostToDevice, c->st));
      HIPCHK(hipMemcpyAsync(c->dBvec.p, net->bVec, sizeof(double) * 3 * ns, hipMemcpyHostToDevice,
                            c->st));
      HIPCHK(hipMemcpyAsync(c->dSegNodes.p, nodes.data(), sizeof(int64_t) * 2 * ns,
                            hipMemcpyHostToDevice, c->st));
    }
    c->dynNodes = nn;
    c->dynConnLd = net->connLd;
    c->dynDrag[0] = drag->edge;
    c->dynDrag[1] = drag->screw;
    c->dynDrag[2] = drag->climb;
    c->dynDrag[3] = drag->line;
    HIPCHK(hipStreamSynchronize(c->st));  // nodes (host vector) is released on return
  });
}

int ddd_net_deriv(ddd_ctx* c, int withPK) {
  return guard(c, "ddd_net_deriv", [&] { dyn_deriv(c, withPK); });
}

int ddd_net_integrate(ddd_ctx* c, const ddd_integ* ip, ddd_itime* t, int withPK) {
  return guard(c, "ddd_net_integrate", [&] {
    if (!ip || !t) bad("NULL argument");
    if (c->dynNodes < 0) bad("no dynamics state (ddd_net_upload_dynamics)");
    const int64_t nn = c->dynNodes;
    if (nn == 0) return;  // IntegrationProcessing.jl:94
    dyn_deriv(c, withPK);
    const int64_t m = 3 * nn;
    HIPCHK(hipMemcpyAsync(c->dCoord0.p, c->dCoord.p, sizeof(double) * m, hipMemcpyDeviceToDevice,
                          c->st));
    HIPCHK(hipMemcpyAsync(c->dVel0.p, c->dVel.p, sizeof(double) * m, hipMemcpyDeviceToDevice,
                          c->st));
    double dt = jmax(t->dt, ip->dtmin);
    const double time = t->time;
    const int64_t step = t->step;
    double dtOld = 0.0;
    bool dtOldGood = false, convergent = false;
    int64_t counter = 1;
    // The reference's loop has no bound; one that keeps failing at a step
    // it once accepted would spin, which a deterministic evaluation cannot.
    for (int64_t iter = 0; !convergent; ++iter) {
      if (iter > 100000) bad("integrate! did not converge in 100000 trials");
      launch_euler(c->dCoord.as<double>(), c->dCoord0.as<double>(), c->dVel0.as<double>(), dt, m,
                   c->st);
      dyn_deriv(c, withPK);
      launch_trapezoid_max(c->dCoord.as<double>(), c->dCoord0.as<double>(), c->dVel.as<double>(),
                           c->dVel0.as<double>(), dt, m,
                           c->dRedPart.as<unsigned long long>(), c->dRed.as<double>(), c->st);
      HIPCHK(hipGetLastError());
      double red[2];
      HIPCHK(hipMemcpyAsync(red, c->dRed.p, sizeof(red), hipMemcpyDeviceToHost, c->st));
      HIPCHK(hipStreamSynchronize(c->st));
      const double maxDist = red[0], maxErr = red[1];
      if (dt <= ip->dtmin) {
        counter = ip->maxiter + 1;
      } else if (maxDist < ip->abstol && maxErr < ip->reltol) {
        dtOld = dt;
        const double factor =
            ip->maxchange *
            std::pow(1.0 / (1.0 + (std::pow(ip->maxchange, ip->exponent) - 1.0) *
                                      (maxErr / ip->reltol)),
                     1.0 / ip->exponent);
        dtOldGood = true;
        dt = jmin(dt * factor, ip->dtmax);
        counter += 1;
      } else if (dtOldGood) {
        dt = dtOld;
        counter = ip->maxiter;
      } else {
        dt = dt / 2.0;
      }
      if (counter > ip->maxiter || dt == ip->dtmax) convergent = true;
    }
    t->dt = dt;
    t->time = time + dt;
    t->step = step + 1;
    check_flags(c);
  });
}

int ddd_net_download_nodes(ddd_ctx* c, double* coord, double* nodeVel, double* nodeForce,
                           int64_t cols) {
  return guard(c, "ddd_net_download_nodes", [&] {
    if (c->dynNodes < 0) bad("no dynamics state (ddd_net_upload_dynamics)");
    const int64_t nn = c->dynNodes;
    if (cols < nn) bad("cols < numNode");
    if (nn == 0) return;
    const size_t bytes = sizeof(double) * 3 * nn;
    if (coord) HIPCHK(hipMemcpyAsync(coord, c->dCoord.p, bytes, hipMemcpyDeviceToHost, c->st));
    if (nodeVel) HIPCHK(hipMemcpyAsync(nodeVel, c->dVel.p, bytes, hipMemcpyDeviceToHost, c->st));
    if (nodeForce)
      HIPCHK(hipMemcpyAsync(nodeForce, c->dNodeF.p, bytes, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

}  // extern "C"
