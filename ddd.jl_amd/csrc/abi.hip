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

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  template <class T> T* as() const { return static_cast<T*>(p); }
  void ensure(size_t need) {
    if (need <= bytes) return;
    const size_t want = std::max(need, bytes + bytes / 2);
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

}  // namespace

struct ddd_farfield {
  GridDev g{};
  int rn = 1;
  int NP = 0;
  int64_t vol = 0;
  bool haveOperator = false;
  DevBuf What, tw, cellOf, counts, cursor, start, tmp, order, M, Mc, Lc, L, counters;
  void release() {
    DevBuf* all[] = {&What, &tw, &cellOf, &counts, &cursor, &start, &tmp, &order, &M, &Mc, &Lc, &L,
                     &counters};
    for (DevBuf* b : all) b->release();
  }
};

struct ddd_ctx {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t st = nullptr;
  std::mutex mu;
  std::string err;
  // segment SoA and pair-kernel scratch
  DevBuf seg, rowPart, colPart, force, rows, out, flags;
  // mesh
  DevBuf meshC, meshCoord, meshConn, uHat;
  // field-point stress
  DevBuf sseg, pts, sigPart, sigma;
  // gather
  DevBuf conn, links, coord, nodeF;
  std::vector<double> h;  // host staging
  DevBuf counter;  // FP64-pair counter of the mixed path
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
void upload_segments(ddd_ctx* c, const ddd_network* net) {
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
  c->seg.ensure(sizeof(double) * F_COUNT * ld);
  HIPCHK(hipMemcpyAsync(c->seg.p, h, sizeof(double) * F_COUNT * ld, hipMemcpyHostToDevice, c->st));
  launch_prep_segments(c->seg.as<double>(), ld, n, c->st);
  HIPCHK(hipGetLastError());
  c->resLd = ld;
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
  r.sinMin = o.sinMin > 0 ? o.sinMin : 1e-3;
  r.K = SigK<float>{float(k.coreRadSq), float(k.mu4pi), float(k.mu4pinu), float(k.nuMu4pinu),
                    float(k.coreRadSq * k.mu8pi)};
  return r;
}

void check_opts(const ddd_opts* o) {
  if (!o) return;
  if (o->mode != DDD_MODE_FP64 && o->mode != DDD_MODE_MIXED && o->mode != DDD_MODE_HYBRID)
    bad("unknown mode");
  if (o->farDist < 0 || o->sinMin < 0) bad("farDist and sinMin must be >= 0");
  if (o->mode == DDD_MODE_HYBRID) {
    if (!o->farfield) bad("hybrid mode needs a farfield (ddd_farfield_create)");
    if (!o->farfield->haveOperator) bad("the farfield has no operator (ddd_farfield_set_operator)");
  }
}

bool is_hybrid(const ddd_opts* o) { return o && o->mode == DDD_MODE_HYBRID; }

// Hybrid pair stage: cells, moments, far-field expansions, near field (adds
// into dev).  The far-field forces are added by hybrid_apply.
void hybrid_pairs(ddd_ctx* c, const ddd_consts& k, int64_t n, double* dev, unsigned* flags,
                  const ddd_opts& o) {
  ddd_farfield* ff = o.farfield;
  const int nc = ff->g.G * ff->g.G * ff->g.G;
  ff->cellOf.ensure(sizeof(int) * std::max<int64_t>(n, 1));
  ff->tmp.ensure(sizeof(int) * std::max<int64_t>(n, 1));
  ff->order.ensure(sizeof(int) * std::max<int64_t>(n, 1));
  ff->counts.ensure(sizeof(unsigned) * nc);
  ff->cursor.ensure(sizeof(unsigned) * nc);
  ff->start.ensure(sizeof(int) * (nc + 1));
  ff->M.ensure(sizeof(float) * kMoments * nc);
  ff->L.ensure(sizeof(float) * kLocal * nc);
  ff->Mc.ensure(sizeof(float2) * kMoments * ff->vol);
  ff->Lc.ensure(sizeof(float2) * kLocal * ff->vol);
  ff->counters.ensure(2 * sizeof(unsigned long long));
  HIPCHK(hipMemsetAsync(ff->counters.p, 0, 2 * sizeof(unsigned long long), c->st));
  const double* seg = c->seg.as<double>();
  launch_cells(seg, c->resLd, (int)n, ff->g, ff->cellOf.as<int>(), ff->counts.as<unsigned>(),
               ff->cursor.as<unsigned>(), ff->start.as<int>(), ff->tmp.as<int>(),
               ff->order.as<int>(), c->st);
  launch_moments(seg, c->resLd, ff->g, ff->start.as<int>(), ff->order.as<int>(), ff->M.as<float>(),
                 c->st);
  launch_far_field(ff->M.as<float>(), ff->g, ff->What.as<float2>(), ff->tw.as<float2>(),
                   ff->Mc.as<float2>(), ff->Lc.as<float2>(), ff->L.as<float>(), c->st);
  const MixedRule mr = mixed_rule(k, o);
  launch_near_rows(seg, c->resLd, ff->g, ff->rn, ff->start.as<int>(), ff->order.as<int>(),
                   seg_consts(k), mr, dev, flags, ff->counters.as<unsigned long long>(), c->st);
  HIPCHK(hipGetLastError());
}

void hybrid_apply(ddd_ctx* c, int64_t n, double* dev, const ddd_opts& o) {
  ddd_farfield* ff = o.farfield;
  launch_far_apply(c->seg.as<double>(), c->resLd, (int)n, ff->g, ff->cellOf.as<int>(),
                   ff->L.as<float>(), dev, c->st);
  HIPCHK(hipGetLastError());
}

// Pair tiles of the whole network (mode from opts) without the reduction.
void segseg_tiles(ddd_ctx* c, const ddd_consts& k, int64_t n, unsigned* flags,
                  const ddd_opts* opts) {
  if (n < 2) return;
  const int64_t nb = (n + kTile - 1) / kTile;
  const int64_t T = tiles_for(nb);
  const size_t part = sizeof(double) * 6 * kTile * (size_t)T;
  c->rowPart.ensure(part);
  c->colPart.ensure(part);
  c->counter.ensure(sizeof(unsigned long long));
  if (opts && opts->mode == DDD_MODE_MIXED) {
    HIPCHK(hipMemsetAsync(c->counter.p, 0, sizeof(unsigned long long), c->st));
    const MixedRule r = mixed_rule(k, *opts);
    launch_segseg_tiles(c->seg.as<double>(), c->resLd, (int)n, seg_consts(k),
                        c->rowPart.as<double>(), c->colPart.as<double>(), flags, c->st, &r,
                        c->counter.as<unsigned long long>());
  } else {
    launch_segseg_tiles(c->seg.as<double>(), c->resLd, (int)n, seg_consts(k),
                        c->rowPart.as<double>(), c->colPart.as<double>(), flags, c->st);
  }
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
  segseg_tiles(c, k, n, flags, opts);
  launch_segseg_reduce(c->rowPart.as<double>(), c->colPart.as<double>(), (int)n, dev, c->st);
  HIPCHK(hipGetLastError());
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
  *out = c;
  return DDD_OK;
}

int ddd_ctx_destroy(ddd_ctx* c) {
  if (!c) return DDD_OK;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->st);
    DevBuf* all[] = {&c->counter, &c->seg, &c->rowPart, &c->colPart, &c->force, &c->rows, &c->out,
                     &c->flags, &c->meshC, &c->meshCoord, &c->meshConn, &c->uHat, &c->sseg,
                     &c->pts, &c->sigPart, &c->sigma, &c->conn, &c->links, &c->coord, &c->nodeF};
    for (DevBuf* b : all) b->release();
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
      launch_segseg_rows(c->seg.as<double>(), c->resLd, (int)n, dr, (int)nIdx, seg_consts(*k), dev,
                         fl, c->st);
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
      launch_segseg_rows(c->seg.as<double>(), c->resLd, (int)n, dr, (int)nIdx, seg_consts(*k), dev,
                         fl, c->st);
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
    c->sseg.ensure(sizeof(double) * 12 * ld);
    HIPCHK(hipMemcpyAsync(c->sseg.p, h, sizeof(double) * 12 * ld, hipMemcpyHostToDevice, c->st));
    launch_prep_sigma_segments(c->sseg.as<double>(), ld, (int)m, c->st);
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
      const SigmaK K{k->coreRadSq, k->mu4pi, k->mu4pinu, k->nuMu4pinu, k->coreRadSq * k->mu8pi};
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
    const int64_t nb = (c->resN + kTile - 1) / kTile;
    const size_t part = sizeof(double) * 6 * kTile * (size_t)tiles_for(nb);
    c->rowPart.ensure(part);
    c->colPart.ensure(part);
    reset_flags(c);
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
    if (is_hybrid(&c->opts)) hybrid_pairs(c, c->resK, n, dev, c->flags.as<unsigned>(), c->opts);
    else segseg_tiles(c, c->resK, n, c->flags.as<unsigned>(), &c->opts);
  } else if (stage == 2) {
    if (is_hybrid(&c->opts)) hybrid_apply(c, n, dev, c->opts);
    else if (n >= 2)
      launch_segseg_reduce(c->rowPart.as<double>(), c->colPart.as<double>(), (int)n, dev, c->st);
  } else {
    bad("stage must be 0, 1 or 2");
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

void fill_stats(ddd_ctx* c, int64_t n, const ddd_opts* o, ddd_stats* st) {
  if (!st) return;
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
    HIPCHK(hipMemcpyAsync(&e, c->counter.p, sizeof(e), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    st->pairsFP64 = e;
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
    std::vector<float2> tw(ff->NP);
    for (int k = 0; k < ff->NP; ++k) {
      const double a = -2.0 * 3.14159265358979323846 * k / ff->NP;
      tw[k] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    try {
      ff->tw.ensure(sizeof(float2) * ff->NP);
      HIPCHK(hipMemcpyAsync(ff->tw.p, tw.data(), sizeof(float2) * ff->NP, hipMemcpyHostToDevice, c->st));
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
      scratch.ensure(sizeof(float2) * kChan * ff->vol);
      ff->What.ensure(sizeof(float2) * kChan * ff->vol);
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

const double* ddd_net_segforce_device(ddd_ctx* c) {
  return (c && c->resN >= 0) ? c->force.as<double>() : nullptr;
}

}  // extern "C"
