/*
 * ddd_hip.h -- C ABI of the MI355X (gfx950) force path for DDD.jl networks.
 *
 * The entry points replace, one for one, the force-path functions of DDD.jl
 * (src/Processing/DislocationForces.jl, src/Processing/DislocationFEM.jl,
 * src/Processing/DislocationMobility.jl).  They take the reference's Julia
 * arrays as they lie in memory -- column-major, Int64 indices 1-based -- so a
 * Julia ccall passes its arrays straight through.
 *
 * Conventions
 *   - All pointers are host pointers unless a name says otherwise.  Every
 *     call is synchronous: results are in the caller's arrays on return and
 *     no host pointer is kept.
 *   - Return value: DDD_OK (0) or a negative DDD_ERR_* code; the message is
 *     in ddd_last_error(ctx).  Nothing aborts or exits.
 *   - `idx`/`nIdx`: NULL/0 means "all numSeg segments"; otherwise nIdx
 *     1-based segment indices (the reference's `idx` argument as an Int64
 *     vector; the Julia shim turns an Int or a range into one).
 *   - Output columns: without DDD_OUT_BY_POSITION, results for segment i go
 *     to column i (the in-place `!` forms); with it, results for idx[k] go to
 *     column k (the out-of-place forms, which return length(idx) columns).
 *   - DDD_ACCUMULATE adds into the output (the reference's in-place forms
 *     accumulate); without it the addressed output columns are overwritten.
 *   - A context is used by one caller thread at a time (calls on one context
 *     are serialised by an internal lock); contexts are independent.
 */
#ifndef DDD_HIP_H
#define DDD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DDD_ABI_VERSION 1

enum {
  DDD_OK = 0,
  DDD_ERR_ARG = -1,       /* bad argument, shape or index            */
  DDD_ERR_OOM = -2,       /* device allocation failed                */
  DDD_ERR_HIP = -3,       /* HIP runtime error                       */
  DDD_ERR_NO_DEVICE = -4, /* no usable gfx950 device                 */
  DDD_ERR_NUMERIC = -5,   /* a numerical guard fired (see message)   */
  DDD_ERR_INTERNAL = -6   /* unexpected C++ exception (see message)  */
};

enum {
  DDD_ACCUMULATE = 1,         /* add into the output instead of overwriting */
  DDD_OUT_BY_POSITION = 2,    /* output column k for idx[k]                 */
  DDD_PK_HALF_TO_NODES = 4    /* ddd_pk_force: 3x2xn, pk/2 to each node     */
};

typedef struct ddd_ctx ddd_ctx;

/* MaterialParameters (src/Type/MaterialType.jl:37-53, derived constants
 * MaterialConstructor.jl:20-29) and DislocationParameters
 * (src/Type/DislocationType.jl:205-255) fields the force path reads. */
typedef struct ddd_consts {
  double mu4pi;     /* mu / (4 pi)                 */
  double mu8pi;     /* mu / (8 pi)                 */
  double mu4pinu;   /* mu / (4 pi (1 - nu))        */
  double nuMu4pinu; /* nu mu / (4 pi (1 - nu))     */
  double omNuInv;   /* 1 / (1 - nu)                */
  double nuOmNuInv; /* nu / (1 - nu)               */
  double coreRad;   /* a                           */
  double coreRadSq; /* a^2                         */
  double coreEnergy;/* Ec                          */
  double sigmaPN;   /* Peierls-Nabarro threshold   */
} ddd_consts;

/* DislocationNetwork (src/Type/DislocationType.jl:324-340) as Julia lays it
 * out.  Sizes are the allocated column counts, used to bounds-check the
 * 1-based indices in segIdx, links and connectivity. */
typedef struct ddd_network {
  int64_t numNode;              /* network.numNode[1]                       */
  int64_t numSeg;               /* network.numSeg[1]                        */
  const int64_t* segIdx;        /* segIdxLd x 3: (link, node1, node2)       */
  int64_t segIdxLd;
  const double* coord;          /* 3 x coordCols                            */
  int64_t coordCols;
  const double* bVec;           /* 3 x bVecCols                             */
  int64_t bVecCols;
  const uint8_t* extSeg;        /* numSeg bytes (Vector{Bool}), may be NULL */
  const int64_t* links;         /* 2 x linksCols (gather only)              */
  int64_t linksCols;
  const int64_t* connectivity;  /* connLd x connCols (gather only)          */
  int64_t connLd;
  int64_t connCols;
} ddd_network;

/* RegularCuboidMesh with linear elements (src/Type/FEMType.jl:212-272) as
 * calc_sigmaHat reads it, plus the dense uHat of ForceDisplacement. */
typedef struct ddd_mesh {
  const double* C;             /* 6 x 6 stiffness, column-major            */
  const double* coord;         /* 3 x numNode                              */
  int64_t numNode;
  const int64_t* connectivity; /* 8 x numElem, 1-based                     */
  int64_t numElem;
  int64_t mx, my, mz;
  double w, h, d;              /* element width, height, depth             */
  const double* uHat;          /* 3 numNode, dense (the shim densifies)    */
} ddd_mesh;

/* Evaluation options (no reference counterpart).  mode DDD_MODE_FP64 is the
 * reference's arithmetic; DDD_MODE_MIXED evaluates well-separated pairs
 * (midpoint distance > farDist and |t1 x t2| > sinMin) in FP32 by
 * Gauss-Legendre quadrature of the non-singular stress, all others in FP64;
 * DDD_MODE_HYBRID adds a learned far field (see ddd_farfield_*). */
enum { DDD_MODE_FP64 = 0, DDD_MODE_MIXED = 1, DDD_MODE_HYBRID = 2 };
typedef struct ddd_farfield ddd_farfield;
typedef struct ddd_opts {
  int mode;
  double farDist;  /* mixed/hybrid: FP32 beyond this midpoint distance (0: 5 coreRad) */
  double sinMin;   /* mixed/hybrid: ... and |t1 x t2| above this (0: 1e-3; < 0: any angle) */
  ddd_farfield* farfield; /* hybrid: cell grid and far-field operator              */
} ddd_opts;

typedef struct ddd_stats {
  uint64_t pairs;       /* pairs evaluated exactly (hybrid: near pairs, each once) */
  uint64_t pairsFP64;   /* of which on the FP64 path                            */
  uint64_t deviceBytes; /* device memory the context's own buffers hold         */
  uint64_t chunks;      /* chunks of the all-pairs triangle (0 for the hybrid)  */
  double tileMs;        /* the all-pairs evaluation's pair-tile kernels alone,
                           summed over chunks (HIP events on the context's
                           stream; plain FP64 and mixed modes, else 0)        */
} ddd_stats;

/* Hybrid far field (no reference counterpart).  A G^3 grid of cubic cells
 * of edge h from corner lo; segments belong to the cell of their midpoint
 * (clamped).  Pairs whose cells are within Chebyshev distance nearRadius are
 * evaluated exactly (mixed rule); all other pairs through 90 cell moments and
 * the operator W: for every cell offset o in [-(G-1), G-1]^3 (x fastest), a
 * 60 x 90 row-major matrix mapping source moments to the 6 x 10 quadratic
 * stress expansion of the target cell.  ddd_farfield_set_operator takes W as
 * (2G-1)^3 x 60 x 90 floats and precomputes its spectrum on the device. */
int ddd_farfield_create(ddd_ctx* ctx, int G, const double lo[3], double h, int nearRadius,
                        ddd_farfield** out);
int ddd_farfield_destroy(ddd_ctx* ctx, ddd_farfield* ff);
int ddd_farfield_set_operator(ddd_ctx* ctx, ddd_farfield* ff, const float* W);
/* The far-field network (no reference counterpart): an MLP from a cell
 * offset to the operator W(o), trained on exact stresses.  Its input
 * features include the core radius in cell units of the grid it was last
 * trained for or installed into (0 after ddd_mlp_create).  Parameters are
 * ddd_mlp_num_params() floats (W1, b1, ..., W4, b4; see csrc/mlp.hpp). */
typedef struct ddd_mlp ddd_mlp;
int ddd_mlp_create(ddd_ctx* ctx, uint64_t seed, ddd_mlp** out);
int ddd_mlp_destroy(ddd_ctx* ctx, ddd_mlp* mlp);
int64_t ddd_mlp_num_params(void);
/* Copy the master parameters out (get != NULL) and/or in (set != NULL). */
int ddd_mlp_params(ddd_ctx* ctx, ddd_mlp* mlp, float* get, const float* set);
/* Loss and gradient on one host batch of ns samples: off3 (3 ns ints, the
 * cell offset source - target), M (90 ns source moments), Lt (60 ns target
 * local expansions, 6 stress components x 10 basis terms), den (ns);
 * loss = sum_s |W(o_s) M_s - Lt_s|^2 / (den_s ns).  For gradient checks. */
int ddd_mlp_loss_grad(ddd_ctx* ctx, ddd_mlp* mlp, int ns, const int* off3, const float* M,
                      const float* Lt, const float* den, float envScale, double* loss,
                      float* grad);

typedef struct ddd_train_opts {
  int64_t samples;     /* training samples (source cell, target point)   */
  int64_t valSamples;  /* held-out samples                                */
  int steps;           /* Adam steps                                      */
  int batch;           /* samples per step                                */
  float lr, lrFinal;   /* learning rate, decayed geometrically            */
  uint64_t seed;
  float denFloor;      /* loss floor, relative to the offset's typical stress (0: 1e-2) */
  float nearFrac;      /* share of sources drawn from the first shells (0: 0.5)          */
  int nearShells;      /* number of those shells outside the near stencil (0: 3)        */
  int plainOffsets;    /* 1: learn W at every offset; 0: at canonical offsets only, the  */
                       /*    rest by cubic symmetry, W(R o) = det(R) T_L W(o) T_M^T      */
} ddd_train_opts;
typedef struct ddd_train_report {
  double lossFirst, lossLast;  /* mean training loss of the first / last 50 steps */
  double relErrVal;            /* held out: sqrt(sum |L_hat - L|^2 / sum |L|^2)   */
  double relErrValNear;        /* the same over samples with Chebyshev offset <= nearRadius + 2 */
  double seconds;
  double relErrShell[8];       /* relErrVal per Chebyshev shell nearRadius + 1 + k (k = 7: and beyond) */
} ddd_train_report;
/* Train mlp on exact local expansions from trainNet binned on ff's grid, then
 * install the operator it predicts into ff (as ddd_farfield_set_operator_mlp).
 * A sample is an occupied source cell and an offset outside the near stencil
 * (a share nearFrac within nearShells shells of it); its target is the L2
 * projection onto the quadratic basis of the source cell's exact FP64 stress
 * at the 27 Gauss points of the target cell.  A network uploaded with
 * ddd_net_upload stays resident. */
int ddd_farfield_train(ddd_ctx* ctx, ddd_farfield* ff, ddd_mlp* mlp, const ddd_consts* k,
                       const ddd_network* trainNet, const ddd_train_opts* opts,
                       ddd_train_report* report);
/* The training targets, for tests: for sample s, the exact local expansion
 * (6 x 10, component-major) of source cell srcCell[s] of net binned on ff's
 * grid over the cell at srcCell[s] - off3[s]. */
int ddd_farfield_local_targets(ddd_ctx* ctx, ddd_farfield* ff, const ddd_consts* k,
                               const ddd_network* net, int ns, const int* srcCell,
                               const int* off3, float* out);
/* Evaluate mlp at every offset of ff's grid and install W(o) (BF16 MFMA). */
int ddd_farfield_set_operator_mlp(ddd_ctx* ctx, ddd_farfield* ff, ddd_mlp* mlp,
                                  const ddd_consts* k);

/* Intermediate results of the last hybrid evaluation, for tests:
 * what 0: cell moments, 90 x G^3 floats; 1: local expansions, 60 x G^3
 * floats; 2: cell of each segment, numSeg int32 (-1 external). */
int ddd_farfield_debug(ddd_ctx* ctx, ddd_farfield* ff, int what, void* out, int64_t bytes);

int ddd_version(void);

/* Self-test (no counterpart in the reference): the all-pairs tile kernel's
 * division, square root, log and atan without range scaling, on given
 * arguments, so that tests can compare them bit for bit with the general
 * IEEE sequences and the host's log/atan.  y[i] = log(x[i]) (which 0),
 * atan(x[i]) (1), sqrt(x[i]) (2), x[2i] / x[2i+1] (3); the ranged forms
 * are exact for the arguments the tile kernel admits (positive normal log
 * arguments, every atan argument, normal operands away from the exponent
 * limits) and give NaN, never a wrong finite value, for zero or infinite
 * log, square-root and division operands. */
int ddd_selftest_ranged(ddd_ctx* ctx, int which, const double* x, int64_t n, double* y);

/* Create a context on HIP device `device`; DDD_ERR_NO_DEVICE when none. */
int ddd_ctx_create(ddd_ctx** ctx, int device);
int ddd_ctx_destroy(ddd_ctx* ctx);
/* Run the context's work on an existing HIP stream (a hipStream_t);
 * NULL returns to the context's own stream. */
int ddd_ctx_set_stream(ddd_ctx* ctx, void* stream);
void* ddd_ctx_stream(ddd_ctx* ctx);
/* Memory budget of one chunk of the all-pairs evaluation (FP64 and mixed
 * modes; no counterpart in the reference): the triangle of pair tiles is
 * evaluated in chunks of whole 64-segment column blocks whose partial sums
 * and near-parallel lists fit in `bytes` (0: the default, 32 GiB), each
 * reduced into segForce in a fixed order before the next. */
int ddd_ctx_set_chunk_bytes(ddd_ctx* ctx, int64_t bytes);
const char* ddd_last_error(ddd_ctx* ctx);

/* calcSegSegForce / calcSegSegForce! (DislocationForces.jl:470-669; pair
 * math :879-1561).  out: 3 x 2 x outCols.  Without idx: all unordered pairs
 * i < j of non-external segments (calcSegSegForceSer!, :778-829); with idx:
 * full rows, forces on the row segment only (calcSegSegForceIdx!,
 * :831-877). */
int ddd_segseg_force(ddd_ctx* ctx, const ddd_consts* k, const ddd_network* net,
                     const int64_t* idx, int64_t nIdx, double* out, int64_t outCols, int flags);

/* ddd_segseg_force over all pairs with evaluation options (opts NULL: plain
 * FP64, identical to ddd_segseg_force without idx); stats may be NULL. */
int ddd_segseg_force_opts(ddd_ctx* ctx, const ddd_consts* k, const ddd_network* net,
                          const ddd_opts* opts, double* out, int64_t outCols, int flags,
                          ddd_stats* stats);

/* calcSelfForce / calcSelfForce! (DislocationForces.jl:320-455).  out:
 * 3 x 2 x outCols; node 1 gets -f, node 2 +f. */
int ddd_self_force(ddd_ctx* ctx, const ddd_consts* k, const ddd_network* net,
                   const int64_t* idx, int64_t nIdx, double* out, int64_t outCols, int flags);

/* calcPKForce / calcPKForce! (DislocationForces.jl:216-300, calc_sigmaHat
 * :92-192).  Without DDD_PK_HALF_TO_NODES out is 3 x outCols (the
 * out-of-place form); with it, 3 x 2 x outCols with pk/2 to each node. */
int ddd_pk_force(ddd_ctx* ctx, const ddd_network* net, const ddd_mesh* mesh,
                 const int64_t* idx, int64_t nIdx, double* out, int64_t outCols, int flags);

/* calc_sigmaHat (DislocationForces.jl:92-192): the FE correction stress of
 * the mesh's dense uHat at P points x0 (3 x P); sigma: 3 x 3 x P, each point's
 * symmetric tensor column-major as the reference's SMatrix{3,3}. */
int ddd_sigma_hat(ddd_ctx* ctx, const ddd_mesh* mesh, const double* x0, int64_t P, double* sigma);

/* calcSegForce / calcSegForce! (DislocationForces.jl:23-71): PK/2 + self +
 * seg-seg.  mesh may be NULL (no FE correction: PK = 0).  out: 3 x 2 x
 * outCols.  The in-place form zeroes the addressed columns first: call it
 * without DDD_ACCUMULATE. */
int ddd_seg_force(ddd_ctx* ctx, const ddd_consts* k, const ddd_network* net,
                  const ddd_mesh* mesh, const int64_t* idx, int64_t nIdx, double* out,
                  int64_t outCols, int flags);

/* calc_sigmaTilde / calc_sigmaTilde! (DislocationFEM.jl:24-303).  x0: 3 x P
 * points; sigma: 6 x P in the order (xx, yy, zz, xy, xz, yz).  External
 * segments are included, as in the reference. */
int ddd_sigma_tilde(ddd_ctx* ctx, const ddd_consts* k, const ddd_network* net,
                    const int64_t* idx, int64_t nIdx, const double* x0, int64_t P,
                    double* sigma, int flags);

/* calc_sigmaTilde without device allocation.  No counterpart in the
 * reference: for callers that cannot call hipMalloc (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  Every
 * array is a DEVICE pointer in the reference's layout: segIdx segIdxLd x 3,
 * coord 3 x coordCols, bVec 3 x bVecCols, idx nIdx 1-based segment indices
 * (NULL: 1..numSeg), x0 3 x P, sigma 6 x P (xx, yy, zz, xy, xz, yz); k is a
 * host pointer.  The call only enqueues work on the context's stream
 * (ddd_ctx_set_stream) and returns; the result is ready when the stream
 * reaches that point.  Temporaries live in a fixed 64 MiB device array per
 * GPU, held until the process ends: kSigmaFields (33) x segments + 6 x
 * points x chunks doubles must fit in 2^23 (calls beyond that return
 * DDD_ERR_ARG before any work).  Calls that overlap on one GPU share that
 * array and race.  A segment whose indices lie outside the arrays gives NaN
 * stresses.  DDD_ACCUMULATE adds into sigma. */
int ddd_sigma_tilde_noalloc(ddd_ctx* ctx, const ddd_consts* k, int64_t numSeg,
                            const int64_t* segIdx, int64_t segIdxLd, const double* coord,
                            int64_t coordCols, const double* bVec, int64_t bVecCols,
                            const int64_t* idx, int64_t nIdx, const double* x0, int64_t P,
                            double* sigma, int flags);

/* Node-force gather of dlnMobility! (DislocationMobility.jl:220-255):
 * nodeForce[:, n] = sum over n's connections of segForce[:, col, link],
 * skipping zero-length links and forces below sigmaPN.  segForce: 3 x 2 x
 * segForceCols; nodeForce: 3 x numNode (overwritten). */
int ddd_node_force(ddd_ctx* ctx, const ddd_consts* k, const ddd_network* net,
                   const double* segForce, int64_t segForceCols, double* nodeForce);

/* DislocationParameters.dragCoeffs (src/Type/DislocationType.jl:151-152,
 * read from the parameter file's "dragCoeffs", src/IO/Input.jl:247-251). */
typedef struct ddd_drag {
  double edge, screw, climb, line;
} ddd_drag;

/* dlnMobility / dlnMobility! for mobBCC (DislocationMobility.jl:16-171 and
 * :186-341): per node, the node force (the gather of ddd_node_force) and the
 * velocity that solves drag * v = f, with the BCC drag tensor summed over the
 * node's links.  idx: nIdx 1-based node indices, NULL for nodes 1..numNode.
 * nodeForce and nodeVel are 3 x outCols.  With DDD_OUT_BY_POSITION the
 * result of idx[k] goes to column k (dlnMobility); otherwise to column
 * idx[k] (dlnMobility!: the other columns are left as they are).  A node
 * whose drag tensor is singular (no links) gets NaN velocities, as the
 * reference's 3x3 solve gives. */
int ddd_mobility_bcc(ddd_ctx* ctx, const ddd_consts* k, const ddd_drag* drag,
                     const ddd_network* net, const double* segForce, int64_t segForceCols,
                     const int64_t* idx, int64_t nIdx, double* nodeForce, double* nodeVel,
                     int64_t outCols, int flags);

/* calc_uTilde / calc_uTilde! (src/Processing/DislocationFEM.jl:355-707):
 * the dislocation displacement at P points (the FE nodes uGammaDln), u:
 * 3 x P, point by point (the reference's out-of-place vector; the in-place
 * form adds value k into dof uDofsDln[k], which the caller does).  Every
 * segment i of links (2 x linksCols, 1-based) with Burgers vector bVec[:, i]
 * is a Barnett triangle closed at the centroid of the mesh vertices; a
 * segment with an external node (label extDln) is closed over its
 * projections onto the mesh surface (solid-angle terms only), and a
 * surface-external link only records the connection, as in the reference.
 * label: nodeTypeDln (Int32) per coord column; vertices 3 x 8, faceNorm and
 * faceMidPt 3 x 6 of the RegularCuboidMesh.  omNuInv8pi = 1/(8 pi (1-nu)),
 * om2NuOmNuInv8pi = (1-2nu)/(8 pi (1-nu)) (MaterialConstructor.jl:28-29).
 * flags: DDD_ACCUMULATE adds into u. */
int ddd_u_tilde(ddd_ctx* ctx, double omNuInv8pi, double om2NuOmNuInv8pi, const ddd_network* net,
                const int32_t* label, const double* vertices, const double* faceNorm,
                const double* faceMidPt, const double* points, int64_t P, double* u, int flags);

/* Result of detectCollision (src/Processing/DislocationTopologyCollision.jl:
 * 1-222): type 0 = :null, 1 = :hinge, 2 = :twoLine; node and link numbers
 * 1-based (0 when there is no collision). */
typedef struct ddd_collision {
  int32_t collision;
  int32_t type;
  int64_t n1s1, n2s1, n1s2, n2s2, s1, s2;
  double L1, L2;
} ddd_collision;

/* detectCollision: hinges (two links at a mobile node), then, if none
 * collides, every pair of links without a shared node, both with mobile
 * ends and not forming a superdislocation; the candidate closer than
 * sqrt(collisionDistSq) and approaching with the largest 1/distSq wins (the
 * first in the reference's loop order on ties).  L1, L2 are those of the
 * last pair evaluated, as the reference returns them.  label: nodeTypeDln
 * (Int32) and nodeVel (3 x coordCols) per coord column; skipSegs: nSkip
 * pairs of link numbers (2 x nSkip, either order). */
int ddd_detect_collision(ddd_ctx* ctx, double collisionDistSq, const ddd_network* net,
                         const int32_t* label, const double* nodeVel, const int64_t* skipSegs,
                         int64_t nSkip, ddd_collision* out);

/* ---- Device-resident path (no counterpart in the reference) ----------
 * For callers that evaluate forces many times on one network (time steps,
 * benchmarks): upload once, compute on the device, fetch when needed.
 * ddd_net_upload packs the network into the context's device buffers;
 * ddd_net_seg_force computes the calcSegForce! total (mesh may be NULL)
 * into the context's device segForce; ddd_net_download copies it into a
 * 3 x 2 x outCols host array (columns 1..numSeg); ddd_net_segforce_device
 * returns the device address of that buffer (6 doubles per segment).
 * A one-shot force call on the same context discards the resident network
 * and its dynamics state (they share the segment buffers). */
int ddd_net_upload(ddd_ctx* ctx, const ddd_consts* k, const ddd_network* net,
                   const ddd_mesh* mesh);
int ddd_net_seg_force(ddd_ctx* ctx, int withPK);
/* One stage of ddd_net_seg_force, for per-kernel timing on the context's
 * stream: 0 = zero segForce + PK/2 + self force, 1 = seg-seg pair tiles,
 * 2 = reduction of the tile partials into segForce; 3 = node forces and
 * velocities from segForce (dlnMobility!, fixed nodes at rest; needs
 * ddd_net_upload_dynamics). */
int ddd_net_stage(ddd_ctx* ctx, int stage, int withPK);
int ddd_net_download(ddd_ctx* ctx, double* out, int64_t outCols);
/* Select the evaluation mode of the resident path (NULL: plain FP64) and
 * read the pair statistics of the last evaluation (synchronises). */
int ddd_net_set_opts(ddd_ctx* ctx, const ddd_opts* opts);
int ddd_net_stats(ddd_ctx* ctx, ddd_stats* stats);
const double* ddd_net_segforce_device(ddd_ctx* ctx);

/* ---- Device-resident dynamics: deriv! and integrate! -------------------
 * IntegrationParameters{AdaptiveEulerTrapezoid} (src/Type/IntegrationType.jl:
 * 28-39) and IntegrationTime (:51-55). */
typedef struct ddd_integ {
  double tmin, tmax, dtmin, dtmax, abstol, reltol, maxchange, exponent;
  int64_t maxiter;
} ddd_integ;
typedef struct ddd_itime {
  double dt, time;
  int64_t step;
} ddd_itime;

/* After ddd_net_upload of the same network, keep its nodes on the device:
 * coord (columns 1..numNode), links, connectivity, bVec, the node labels
 * (nodeTypeDln, Int32 as Julia stores the enum) and the drag coefficients.
 * Needs the reference's segment numbering segIdx[i, 1] == i
 * (_getSegmentIdx!, src/Type/DislocationUtil.jl:283-298).  Surface-mobile
 * nodes (srfMobDln) are refused: the reference's surface correction
 * (IntegrationProcessing.jl:34-60) indexes label with a tuple, which Julia
 * rejects, so no reference run with such nodes gets past it. */
int ddd_net_upload_dynamics(ddd_ctx* ctx, const ddd_network* net, const int32_t* label,
                            const ddd_drag* drag);
/* deriv! (src/Processing/IntegrationProcessing.jl:13-62) on the device:
 * calcSegForce! at the device coordinates (in the mode of ddd_net_set_opts;
 * withPK adds the uploaded mesh's PK force), dlnMobility! for mobBCC, and
 * zero velocity for fixed and external nodes. */
int ddd_net_deriv(ddd_ctx* ctx, int withPK);
/* integrate! for AdaptiveEulerTrapezoid (IntegrationProcessing.jl:81-168):
 * moves the device coordinates and returns in *t the (dt, time, step) the
 * reference returns.  Each trial runs on the device; only the two error
 * maxima come back to the host. */
int ddd_net_integrate(ddd_ctx* ctx, const ddd_integ* ip, ddd_itime* t, int withPK);
/* Copy the device coord, nodeVel and nodeForce into 3 x cols host arrays
 * (columns 1..numNode; any pointer may be NULL). */
int ddd_net_download_nodes(ddd_ctx* ctx, double* coord, double* nodeVel, double* nodeForce,
                           int64_t cols);

#ifdef __cplusplus
}
#endif
#endif /* DDD_HIP_H */
