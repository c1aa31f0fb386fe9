"""The reference's force-path API, routed to the HIP library.

Names, argument order, output shapes and in-place semantics follow
src/Processing/DislocationForces.jl, src/Processing/DislocationFEM.jl and the
force half of src/Processing/DislocationMobility.jl.  Julia's ``f!`` is
spelled ``f_`` here.  ``idx`` is None, an int, a range or a sequence of
1-based segment indices, as in the reference.

Every function calls libdddhip.so; there is no CPU path behind them.
"""
from __future__ import annotations

import ctypes as _c

import numpy as np

from . import _lib as L

C_void = _c.c_void_p
_byref = _c.byref


def _idx(idx):
    if idx is None:
        return None
    if isinstance(idx, (int, np.integer)):
        return np.array([int(idx)], dtype=np.int64)
    return np.ascontiguousarray(np.asarray(list(idx) if isinstance(idx, range) else idx,
                                           dtype=np.int64).ravel())


def _ctx(ctx):
    return ctx if ctx is not None else L.default_context()


def _call(ctx, fn, *args):
    ctx.check(getattr(L.lib(), fn)(ctx.h, *args))


def _nidx(ix):
    return (L.ptr(ix, "i"), len(ix)) if ix is not None else (None, 0)


def calcSegSegForce(dlnParams, matParams, network, idx=None, ctx=None):
    """calcSegSegForce (DislocationForces.jl:470-548): 3 x 2 x numSeg, or
    3 x 2 x length(idx) with forces on the idx rows only."""
    ctx = _ctx(ctx)
    nv = L.NetView(network)
    k = L.consts(dlnParams, matParams)
    ix = _idx(idx)
    m = int(network.numSeg[0]) if ix is None else len(ix)
    out = np.zeros((3, 2, m), order="F")
    _call(ctx, "ddd_segseg_force", k, nv.s, *_nidx(ix), L.ptr(out, "d"), m,
          L.DDD_OUT_BY_POSITION)
    return out


def setChunkBytes(nbytes, ctx=None):
    """Memory budget of one chunk of the all-pairs evaluation (0: 32 GiB)."""
    ctx = _ctx(ctx)
    _call(ctx, "ddd_ctx_set_chunk_bytes", int(nbytes))


def calcSegSegForceMixed(dlnParams, matParams, network, farDist=0.0, sinMin=0.0, ctx=None):
    """All-pairs seg-seg force with the mixed-precision rule (no reference
    counterpart): pairs farther than farDist (default 5 coreRad) with
    |t1 x t2| > sinMin (default 1e-3) in FP32 by quadrature, others in FP64.
    Returns (3 x 2 x numSeg forces, pairs on the FP64 path)."""
    ctx = _ctx(ctx)
    nv = L.NetView(network)
    k = L.consts(dlnParams, matParams)
    n = int(network.numSeg[0])
    out = np.zeros((3, 2, n), order="F")
    o = L.ddd_opts(mode=L.DDD_MODE_MIXED, farDist=farDist, sinMin=sinMin)
    st = L.ddd_stats()
    _call(ctx, "ddd_segseg_force_opts", k, nv.s, o, L.ptr(out, "d"), n, 0, st)
    return out, int(st.pairsFP64)


def calcSegSegForce_(dlnParams, matParams, network, idx=None, ctx=None):
    """calcSegSegForce! (DislocationForces.jl:563-669): adds into
    network.segForce."""
    ctx = _ctx(ctx)
    nv = L.NetView(network)
    k = L.consts(dlnParams, matParams)
    ix = _idx(idx)
    sf = network.segForce
    assert sf.flags.f_contiguous and sf.dtype == np.float64
    _call(ctx, "ddd_segseg_force", k, nv.s, *_nidx(ix), L.ptr(sf, "d"), sf.shape[2],
          L.DDD_ACCUMULATE)


def calcSelfForce(dlnParams, matParams, network, idx=None, ctx=None):
    """calcSelfForce (DislocationForces.jl:320-352): (node1, node2), each
    3 x numSeg (or 3 x length(idx))."""
    ctx = _ctx(ctx)
    nv = L.NetView(network)
    k = L.consts(dlnParams, matParams)
    ix = _idx(idx)
    m = int(network.numSeg[0]) if ix is None else len(ix)
    out = np.zeros((3, 2, m), order="F")
    _call(ctx, "ddd_self_force", k, nv.s, *_nidx(ix), L.ptr(out, "d"), m, L.DDD_OUT_BY_POSITION)
    return out[:, 0, :].copy(order="F"), out[:, 1, :].copy(order="F")


def calcSelfForce_(dlnParams, matParams, network, idx=None, ctx=None):
    """calcSelfForce! (DislocationForces.jl:354-388)."""
    ctx = _ctx(ctx)
    nv = L.NetView(network)
    k = L.consts(dlnParams, matParams)
    ix = _idx(idx)
    sf = network.segForce
    _call(ctx, "ddd_self_force", k, nv.s, *_nidx(ix), L.ptr(sf, "d"), sf.shape[2],
          L.DDD_ACCUMULATE)


def calcPKForce(mesh, forceDisplacement, network, idx=None, ctx=None):
    """calcPKForce (DislocationForces.jl:216-243): 3 x numSeg (or length(idx))."""
    ctx = _ctx(ctx)
    nv = L.NetView(network)
    mv = L.MeshView(mesh, forceDisplacement)
    ix = _idx(idx)
    m = int(network.numSeg[0]) if ix is None else len(ix)
    out = np.zeros((3, m), order="F")
    _call(ctx, "ddd_pk_force", nv.s, mv.s, *_nidx(ix), L.ptr(out, "d"), m, L.DDD_OUT_BY_POSITION)
    return out


def calcPKForce_(mesh, forceDisplacement, network, idx=None, ctx=None):
    """calcPKForce! (DislocationForces.jl:244-273): pk/2 added to both nodes."""
    ctx = _ctx(ctx)
    nv = L.NetView(network)
    mv = L.MeshView(mesh, forceDisplacement)
    ix = _idx(idx)
    sf = network.segForce
    _call(ctx, "ddd_pk_force", nv.s, mv.s, *_nidx(ix), L.ptr(sf, "d"), sf.shape[2],
          L.DDD_ACCUMULATE | L.DDD_PK_HALF_TO_NODES)


def calc_σHat(mesh, forceDisplacement, x0, ctx=None):
    """calc_σHat (DislocationForces.jl:92-192): the 3 x 3 stress of the FE
    correction at point x0; for a 3 x P array of points, 3 x 3 x P."""
    ctx = _ctx(ctx)
    mv = L.MeshView(mesh, forceDisplacement)
    x = np.asarray(x0, dtype=np.float64)
    single = x.ndim == 1
    pts = _points(x)
    P = pts.shape[1]
    out = np.zeros((3, 3, P), order="F")
    _call(ctx, "ddd_sigma_hat", mv.s, L.ptr(pts, "d"), P, L.ptr(out, "d"))
    return out[:, :, 0].copy(order="F") if single else out


def selftest_ranged(which, x, ctx=None):
    """The all-pairs tile kernel's ranged log (which 0), atan (1), sqrt (2)
    of every x, or its division x[0::2] / x[1::2] (3), evaluated on the GPU
    (ddd_selftest_ranged; a self-test, no reference counterpart)."""
    ctx = _ctx(ctx)
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.size // 2 if which == 3 else x.size
    y = np.empty(n)
    _call(ctx, "ddd_selftest_ranged", int(which), L.ptr(x, "d"), n, L.ptr(y, "d"))
    return y


def calcSegForce(dlnParams, matParams, mesh, forceDisplacement, network, idx=None, ctx=None):
    """calcSegForce (DislocationForces.jl:23-46): seg-seg + self + PK/2."""
    ctx = _ctx(ctx)
    nv = L.NetView(network)
    mv = L.MeshView(mesh, forceDisplacement) if mesh is not None else None
    k = L.consts(dlnParams, matParams)
    ix = _idx(idx)
    m = int(network.numSeg[0]) if ix is None else len(ix)
    out = np.zeros((3, 2, m), order="F")
    _call(ctx, "ddd_seg_force", k, nv.s, mv.s if mv else None, *_nidx(ix), L.ptr(out, "d"), m,
          L.DDD_OUT_BY_POSITION)
    return out


def calcSegForce_(dlnParams, matParams, mesh, forceDisplacement, network, idx=None, ctx=None):
    """calcSegForce! (DislocationForces.jl:48-71): zero segForce over the
    range, then add PK/2, self and seg-seg."""
    ctx = _ctx(ctx)
    nv = L.NetView(network)
    mv = L.MeshView(mesh, forceDisplacement) if mesh is not None else None
    k = L.consts(dlnParams, matParams)
    ix = _idx(idx)
    sf = network.segForce
    if ix is None:
        sf[:, :, :int(network.numSeg[0])] = 0
    else:
        sf[:, :, ix - 1] = 0
    _call(ctx, "ddd_seg_force", k, nv.s, mv.s if mv else None, *_nidx(ix), L.ptr(sf, "d"),
          sf.shape[2], L.DDD_ACCUMULATE)


def _points(x0):
    x = np.asarray(x0, dtype=np.float64)
    if x.ndim == 1:
        x = x.reshape(3, 1)
    return np.asfortranarray(x.reshape(3, -1, order="F"))


def calc_σTilde(x0, dlnParams, matParams, network, idx=None, ctx=None):
    """calc_σTilde (DislocationFEM.jl:24-41): 6 x P stress (xx, yy, zz, xy,
    xz, yz), or a 6-vector for a single point."""
    x = _points(x0)
    σ = np.zeros((6, x.shape[1]), order="F")
    calc_σTilde_(σ, x, dlnParams, matParams, network, idx, ctx=ctx)
    return σ[:, 0].copy() if x.shape[1] == 1 else σ


def calc_σTilde_(σ, x0, dlnParams, matParams, network, idx=None, ctx=None):
    """calc_σTilde! (DislocationFEM.jl:67-84): overwrites σ (6 x P)."""
    ctx = _ctx(ctx)
    x = _points(x0)
    assert σ.shape[0] == 6 and σ.size == 6 * x.shape[1], "calc_σTilde!: σ must be 6 x P"
    assert σ.dtype == np.float64 and (σ.flags.f_contiguous or σ.ndim == 1)
    nv = L.NetView(network)
    k = L.consts(dlnParams, matParams)
    ix = _idx(idx)
    _call(ctx, "ddd_sigma_tilde", k, nv.s, *_nidx(ix), L.ptr(x, "d"), x.shape[1],
          L.ptr(σ, "d"), 0)
    return σ


def calcNodeForce_(dlnParams, matParams, network, ctx=None):
    """The node-force gather of dlnMobility! (DislocationMobility.jl:220-255,
    336): network.nodeForce[:, 1:numNode] from network.segForce."""
    ctx = _ctx(ctx)
    nv = L.NetView(network)
    k = L.consts(dlnParams, matParams)
    sf = np.asfortranarray(network.segForce)
    n = int(network.numNode[0])
    nf = np.zeros((3, n), order="F")
    _call(ctx, "ddd_node_force", k, nv.s, L.ptr(sf, "d"), sf.shape[2], L.ptr(nf, "d"))
    network.nodeForce[:, :n] = nf
    return nf


def u_tilde_points(matParams, network, mesh, points, ctx=None):
    """Displacement of the network's dislocations at the points (3 x P):
    the Barnett-triangle sum of _calc_uTilde! (DislocationFEM.jl:381-600)
    before it is scattered into dofs; column i belongs to point i."""
    ctx = _ctx(ctx)
    x = np.asfortranarray(np.asarray(points, dtype=np.float64).reshape(3, -1, order="F"))
    P = x.shape[1]
    out = np.zeros((3, P), order="F")
    nv = L.NetView(network)
    lab = np.ascontiguousarray(network.label, dtype=np.int32)
    if len(lab) < nv.coord.shape[1]:
        raise ValueError("network.label needs one entry per coord column")
    keep = [np.asfortranarray(a, dtype=np.float64)
            for a in (mesh.vertices, mesh.faceNorm, mesh.faceMidPt)]
    _call(ctx, "ddd_u_tilde", float(matParams.omνInv8π), float(matParams.om2νomνInv8π), nv.s,
          lab.ctypes.data_as(_c.POINTER(_c.c_int32)), *(L.ptr(a, "d") for a in keep),
          L.ptr(x, "d"), P, L.ptr(out, "d"), 0)
    return out


def calc_uTilde(mesh, boundaries, matParams, network, ctx=None):
    """calc_uTilde (DislocationFEM.jl:369-379): displacements at the nodes
    boundaries.uGammaDln, interleaved (x, y, z of each node in turn)."""
    pts = mesh.coord[:, np.asarray(boundaries.uGammaDln) - 1]
    return u_tilde_points(matParams, network, mesh, pts, ctx).ravel(order="F")


def calc_uTilde_(forceDisplacement, mesh, boundaries, matParams, network, ctx=None):
    """calc_uTilde! (DislocationFEM.jl:355-367): zero forceDisplacement.uTilde
    and add value k of calc_uTilde into dof boundaries.uDofsDln[k] (the
    reference's mapping; its dof list runs x, then y, then z)."""
    v = calc_uTilde(mesh, boundaries, matParams, network, ctx)
    u = forceDisplacement.uTilde
    u[:] = 0.0
    np.add.at(u, np.asarray(boundaries.uDofsDln) - 1, v)
    return u


_COLLISION_TYPES = {0: "null", 1: "hinge", 2: "twoLine"}


def detectCollision(dlnParams, network, skipSegs=(), ctx=None):
    """detectCollision (DislocationTopologyCollision.jl:1-222): (collision,
    type, n1s1, n2s1, n1s2, n2s2, s1, s2, L1, L2) with type "null", "hinge"
    or "twoLine"; skipSegs: pairs of link numbers to ignore."""
    ctx = _ctx(ctx)
    nv = L.NetView(network)
    lab = np.ascontiguousarray(network.label, dtype=np.int32)
    vel = np.asfortranarray(network.nodeVel, dtype=np.float64)
    if len(lab) < nv.coord.shape[1] or vel.shape[1] < nv.coord.shape[1]:
        raise ValueError("network.label and nodeVel need one entry per coord column")
    sk = np.ascontiguousarray(np.asarray(list(skipSegs), dtype=np.int64).reshape(-1, 2))
    res = L.ddd_collision()
    _call(ctx, "ddd_detect_collision", float(dlnParams.collisionDistSq), nv.s,
          lab.ctypes.data_as(_c.POINTER(_c.c_int32)), L.ptr(vel, "d"),
          L.ptr(sk, "i") if len(sk) else None, len(sk), _byref(res))
    return (bool(res.collision), _COLLISION_TYPES[res.type], res.n1s1, res.n2s1, res.n1s2,
            res.n2s2, res.s1, res.s2, res.L1, res.L2)


def _mobility(dlnParams, matParams, network, ix, nf, nv_out, cols, flags, ctx):
    if dlnParams.mobility != "mobBCC":
        raise NotImplementedError(f"mobility law {dlnParams.mobility!r}: only mobBCC exists")
    nv = L.NetView(network)
    k = L.consts(dlnParams, matParams)
    dr = L.drag(dlnParams)
    sf = np.asfortranarray(network.segForce, dtype=np.float64)
    _call(ctx, "ddd_mobility_bcc", k, _byref(dr), nv.s, L.ptr(sf, "d"), sf.shape[2], *_nidx(ix),
          L.ptr(nf, "d"), L.ptr(nv_out, "d"), cols, flags)


def dlnMobility(dlnParams, matParams, network, idx=None, ctx=None):
    """dlnMobility for mobBCC (DislocationMobility.jl:16-171): (nodeForce,
    nodeVel), each 3 x numNode, or 3 x length(idx) with column k for the
    k-th node of idx."""
    ctx = _ctx(ctx)
    ix = _idx(idx)
    n = int(network.numNode[0]) if ix is None else len(ix)
    nf = np.zeros((3, n), order="F")
    vel = np.zeros((3, n), order="F")
    _mobility(dlnParams, matParams, network, ix, nf, vel, n, L.DDD_OUT_BY_POSITION, ctx)
    return nf, vel


def dlnMobility_(dlnParams, matParams, network, idx=None, ctx=None):
    """dlnMobility! for mobBCC (DislocationMobility.jl:186-341): writes
    network.nodeForce[:, node] and network.nodeVel[:, node] for the nodes of
    idx (all numNode nodes by default) and returns the network."""
    ctx = _ctx(ctx)
    ix = _idx(idx)
    for name in ("nodeForce", "nodeVel"):
        a = getattr(network, name)
        if not (a.flags.f_contiguous and a.dtype == np.float64 and a.shape[0] == 3):
            raise TypeError(f"network.{name} must be a 3 x nodeBuf Float64 column-major array")
    cols = min(network.nodeForce.shape[1], network.nodeVel.shape[1])
    _mobility(dlnParams, matParams, network, ix, network.nodeForce, network.nodeVel, cols, 0, ctx)
    return network


class FarField:
    """Hybrid far field (ddd_farfield_*): a G^3 cell grid of edge h from
    corner lo, near radius rn (in cells), and the operator W
    ((2G-1)^3 x 60 x 90) mapping source-cell moments to target-cell stress
    expansions."""

    def __init__(self, G, lo, h, nearRadius, ctx=None):
        self.ctx = _ctx(ctx)
        self.G, self.h, self.rn = int(G), float(h), int(nearRadius)
        self.lo = np.ascontiguousarray(lo, dtype=np.float64)
        self.h_ptr = C_void()
        _call(self.ctx, "ddd_farfield_create", self.G, L.ptr(self.lo, "d"), self.h, self.rn,
              _byref(self.h_ptr))

    def set_operator(self, W):
        S = 2 * self.G - 1
        W = np.ascontiguousarray(W, dtype=np.float32).reshape(S ** 3, 60, 90)
        _call(self.ctx, "ddd_farfield_set_operator", self.h_ptr,
              W.ctypes.data_as(_c.POINTER(_c.c_float)))

    def debug(self, what, numSeg=0):
        nc = self.G ** 3
        if what == 0:
            out = np.zeros((90, nc), dtype=np.float32)
        elif what == 1:
            out = np.zeros((60, nc), dtype=np.float32)
        else:
            out = np.zeros(numSeg, dtype=np.int32)
        _call(self.ctx, "ddd_farfield_debug", self.h_ptr, int(what),
              out.ctypes.data_as(_c.c_void_p), out.nbytes)
        return out

    def close(self):
        if self.h_ptr:
            L.lib().ddd_farfield_destroy(self.ctx.h, self.h_ptr)
            self.h_ptr = C_void()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FarFieldMLP:
    """The far-field network (ddd_mlp_*): offset features -> 256 -> 256 ->
    256 -> 5408, SiLU, BF16 MFMA products with FP32 accumulation; FP32 master
    parameters and Adam state on the device."""

    def __init__(self, seed=0, ctx=None):
        self.ctx = _ctx(ctx)
        self.h_ptr = C_void()
        _call(self.ctx, "ddd_mlp_create", int(seed), _byref(self.h_ptr))

    @staticmethod
    def num_params():
        return int(L.lib().ddd_mlp_num_params())

    def get_params(self):
        out = np.zeros(self.num_params(), dtype=np.float32)
        _call(self.ctx, "ddd_mlp_params", self.h_ptr, out.ctypes.data_as(_c.POINTER(_c.c_float)),
              None)
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.num_params():
            raise ValueError("wrong parameter count")
        _call(self.ctx, "ddd_mlp_params", self.h_ptr, None,
              p.ctypes.data_as(_c.POINTER(_c.c_float)))

    def loss_grad(self, off3, M, Lt, den, envScale):
        """Loss and parameter gradient on one batch of (offset, source
        moments M (90), target local expansion Lt (6 x 10)) samples
        (see ddd_mlp_loss_grad)."""
        f = _c.POINTER(_c.c_float)
        off3 = np.ascontiguousarray(off3, dtype=np.int32)
        ns = off3.shape[0]
        M = np.ascontiguousarray(M, dtype=np.float32).reshape(ns, 90)
        Lt = np.ascontiguousarray(Lt, dtype=np.float32).reshape(ns, 60)
        den = np.ascontiguousarray(den, dtype=np.float32).reshape(ns)
        loss = _c.c_double()
        grad = np.zeros(self.num_params(), dtype=np.float32)
        _call(self.ctx, "ddd_mlp_loss_grad", self.h_ptr, int(ns),
              off3.ctypes.data_as(_c.POINTER(_c.c_int)), M.ctypes.data_as(f),
              Lt.ctypes.data_as(f), den.ctypes.data_as(f), float(envScale), _byref(loss),
              grad.ctypes.data_as(f))
        return loss.value, grad

    def close(self):
        if self.h_ptr:
            L.lib().ddd_mlp_destroy(self.ctx.h, self.h_ptr)
            self.h_ptr = C_void()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def trainFarField(dlnParams, matParams, farfield, mlp, trainNetwork, samples=1 << 20,
                  valSamples=1 << 15, steps=3000, batch=2048, lr=1e-3, lrFinal=3e-5, seed=1,
                  denFloor=0.0, nearFrac=0.0, nearShells=0, cubic=True):
    """Train mlp on exact stresses of trainNetwork binned on farfield's grid
    and install the operator it predicts.  cubic: learn the operator at the
    canonical offsets only and expand it by the cubic symmetry of isotropic
    elasticity (otherwise at every offset).  Returns the report as a dict."""
    ctx = farfield.ctx
    nv = L.NetView(trainNetwork)
    k = L.consts(dlnParams, matParams)
    o = L.ddd_train_opts(samples=int(samples), valSamples=int(valSamples), steps=int(steps),
                         batch=int(batch), lr=float(lr), lrFinal=float(lrFinal), seed=int(seed),
                         denFloor=float(denFloor), nearFrac=float(nearFrac),
                         nearShells=int(nearShells), plainOffsets=0 if cubic else 1)
    r = L.ddd_train_report()
    _call(ctx, "ddd_farfield_train", farfield.h_ptr, mlp.h_ptr, k, nv.s, o, r)
    out = {f: getattr(r, f) for f, _ in L.ddd_train_report._fields_}
    out["relErrShell"] = list(r.relErrShell)
    return out


def farFieldLocalTargets(dlnParams, matParams, farfield, network, srcCell, off3):
    """Exact local expansions (ns x 60) of source cells over the cells at
    srcCell - off3, from network binned on farfield's grid (test hook)."""
    nv = L.NetView(network)
    k = L.consts(dlnParams, matParams)
    src = np.ascontiguousarray(srcCell, dtype=np.int32)
    off = np.ascontiguousarray(off3, dtype=np.int32).reshape(-1, 3)
    out = np.zeros((len(src), 60), dtype=np.float32)
    _call(farfield.ctx, "ddd_farfield_local_targets", farfield.h_ptr, k, nv.s, len(src),
          src.ctypes.data_as(_c.POINTER(_c.c_int)), off.ctypes.data_as(_c.POINTER(_c.c_int)),
          out.ctypes.data_as(_c.POINTER(_c.c_float)))
    return out


def setFarFieldFromMLP(dlnParams, matParams, farfield, mlp):
    """Install the operator mlp predicts at every offset of farfield's grid."""
    k = L.consts(dlnParams, matParams)
    _call(farfield.ctx, "ddd_farfield_set_operator_mlp", farfield.h_ptr, mlp.h_ptr, k)


def calcSegSegForceHybrid(dlnParams, matParams, network, farfield, farDist=0.0, sinMin=0.0,
                          ctx=None):
    """All-pairs seg-seg force, hybrid: near pairs exact (mixed rule), far
    pairs through the far-field operator.  Returns (3 x 2 x numSeg forces,
    near pairs, each counted once, of which FP64)."""
    ctx = _ctx(ctx)
    nv = L.NetView(network)
    k = L.consts(dlnParams, matParams)
    n = int(network.numSeg[0])
    out = np.zeros((3, 2, n), order="F")
    o = L.ddd_opts(mode=L.DDD_MODE_HYBRID, farDist=farDist, sinMin=sinMin,
                   farfield=farfield.h_ptr)
    st = L.ddd_stats()
    _call(ctx, "ddd_segseg_force_opts", k, nv.s, o, L.ptr(out, "d"), n, 0, st)
    return out, int(st.pairs), int(st.pairsFP64)


class ResidentNetwork:
    """Device-resident evaluation (ddd_net_*): upload once, evaluate many
    times.  Used by time-stepping callers and the benchmark."""

    def __init__(self, dlnParams, matParams, network, mesh=None, forceDisplacement=None,
                 ctx=None):
        self.ctx = _ctx(ctx)
        self.nv = L.NetView(network)
        self.mv = L.MeshView(mesh, forceDisplacement) if mesh is not None else None
        self.k = L.consts(dlnParams, matParams)
        self.numSeg = int(network.numSeg[0])
        _call(self.ctx, "ddd_net_upload", self.k, self.nv.s, self.mv.s if self.mv else None)

    def seg_force(self, with_pk=False):
        _call(self.ctx, "ddd_net_seg_force", 1 if with_pk else 0)

    def reupload(self):
        """Upload the network again from the host arrays it was built from
        (discards device coordinates moved by integrate and the dynamics
        state)."""
        _call(self.ctx, "ddd_net_upload", self.k, self.nv.s, self.mv.s if self.mv else None)

    def set_mode(self, mode="fp64", farDist=0.0, sinMin=0.0, farfield=None):
        m = {"fp64": L.DDD_MODE_FP64, "mixed": L.DDD_MODE_MIXED, "hybrid": L.DDD_MODE_HYBRID}[mode]
        _call(self.ctx, "ddd_net_set_opts",
              L.ddd_opts(mode=m, farDist=farDist, sinMin=sinMin,
                         farfield=farfield.h_ptr if farfield is not None else None))

    def stats(self):
        st = L.ddd_stats()
        _call(self.ctx, "ddd_net_stats", st)
        return int(st.pairs), int(st.pairsFP64)

    def footprint(self):
        """(device bytes the context holds, chunks of the last all-pairs
        evaluation)."""
        st = L.ddd_stats()
        _call(self.ctx, "ddd_net_stats", st)
        return int(st.deviceBytes), int(st.chunks)

    def tile_kernel_ms(self):
        """Time of the last evaluation's pair-tile kernels alone (plain FP64
        and mixed modes), from HIP events the library records around them."""
        st = L.ddd_stats()
        _call(self.ctx, "ddd_net_stats", st)
        return float(st.tileMs)

    def stage(self, stage, with_pk=False):
        _call(self.ctx, "ddd_net_stage", int(stage), 1 if with_pk else 0)

    def download(self):
        out = np.zeros((3, 2, self.numSeg), order="F")
        _call(self.ctx, "ddd_net_download", L.ptr(out, "d"), self.numSeg)
        return out

    def stream(self):
        return L.lib().ddd_ctx_stream(self.ctx.h)

    # ---- dynamics: deriv! and integrate! with the nodes on the device
    def upload_dynamics(self, dlnParams, network):
        """Keep coord, links, connectivity, bVec and labels on the device
        (ddd_net_upload_dynamics) for deriv/integrate."""
        self.numNode = int(network.numNode[0])
        self._label = np.ascontiguousarray(network.label, dtype=np.int32)
        self._drag = L.drag(dlnParams)
        _call(self.ctx, "ddd_net_upload_dynamics", self.nv.s,
              self._label.ctypes.data_as(_c.POINTER(_c.c_int32)), _byref(self._drag))

    def deriv(self, with_pk=False):
        _call(self.ctx, "ddd_net_deriv", 1 if with_pk else 0)

    def integrate(self, intParams, intVars, with_pk=False):
        """One integrate! step; returns the new IntegrationTime."""
        from .types import IntegrationTime
        t = L.ddd_itime(float(intVars.dt), float(intVars.time), int(intVars.step))
        ip = L.integ(intParams)
        _call(self.ctx, "ddd_net_integrate", _byref(ip), _byref(t), 1 if with_pk else 0)
        return IntegrationTime(t.dt, t.time, t.step)

    def download_nodes(self):
        """(coord, nodeVel, nodeForce), each 3 x numNode."""
        out = [np.zeros((3, self.numNode), order="F") for _ in range(3)]
        _call(self.ctx, "ddd_net_download_nodes", *(L.ptr(a, "d") for a in out), self.numNode)
        return tuple(out)


def _dynamics(dlnParams, matParams, mesh, forceDisplacement, network, ctx):
    if dlnParams.mobility != "mobBCC":
        raise NotImplementedError(f"mobility law {dlnParams.mobility!r}: only mobBCC exists")
    r = ResidentNetwork(dlnParams, matParams, network, mesh, forceDisplacement, ctx)
    r.upload_dynamics(dlnParams, network)
    return r


def _store(r, network):
    n, ns = r.numNode, r.numSeg
    c, v, f = r.download_nodes()
    network.coord[:, :n] = c
    network.nodeVel[:, :n] = v
    network.nodeForce[:, :n] = f
    network.segForce[:, :, :ns] = r.download()


def deriv_(dlnParams, matParams, mesh, forceDisplacement, network, ctx=None):
    """deriv! (IntegrationProcessing.jl:13-62): calcSegForce!, dlnMobility!
    and zero velocity of fixed nodes, on the device; fills network.segForce,
    nodeForce and nodeVel.  mesh None leaves out the PK force."""
    r = _dynamics(dlnParams, matParams, mesh, forceDisplacement, network, _ctx(ctx))
    r.deriv(with_pk=mesh is not None)
    _store(r, network)


def integrate_(intParams, intVars, dlnParams, matParams, mesh, forceDisplacement, network,
               ctx=None):
    """integrate! for AdaptiveEulerTrapezoid (IntegrationProcessing.jl:81-168)
    on the device: moves network.coord, fills segForce, nodeForce, nodeVel
    and returns the new IntegrationTime.  For many steps keep the network
    resident (ResidentNetwork.upload_dynamics / integrate) instead."""
    r = _dynamics(dlnParams, matParams, mesh, forceDisplacement, network, _ctx(ctx))
    t = r.integrate(intParams, intVars, with_pk=mesh is not None)
    _store(r, network)
    return t
