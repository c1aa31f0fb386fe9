"""ctypes front end of the CPU oracle.  TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module, and only as the checker; the product path
(the HIP library and its Python mirror) never does.

Every function takes the same host structs as the product API
(``dddhip.types``) and returns arrays in the reference's layout.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libddd_oracle.so"
_lib = None

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int64)
up = C.POINTER(C.c_uint8)


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        i64, d, i32 = C.c_int64, C.c_double, C.c_int
        mesh = [dp, dp, ip, i64, i64, i64, d, d, d, dp]
        net = [i64, ip, i64, dp, dp, up]
        sig = {
            "orc_segseg_pair": (i32, [i32, dp] + [dp] * 7),
            "orc_count_segseg_pair": (i32, [dp] * 8),
            "orc_count_sigma_tilde": (None, [dp] * 6),
            "orc_count_self_force": (None, [dp] * 4),
            "orc_segseg_ser": (None, [i32, dp] + net + [dp]),
            "orc_segseg_par_faithful": (None, [dp] + net + [dp, i32]),
            "orc_segseg_idx": (None, [dp] + net + [ip, i64, dp, i32]),
            "orc_self_force": (None, [dp] + net + [ip, i64, dp, i32]),
            "orc_sigma_hat": (None, mesh + [dp, dp]),
            "orc_pk_force": (None, mesh + net + [ip, i64, dp, i32, i32]),
            "orc_sigma_tilde": (None, [dp, i64, ip, i64, dp, dp, ip, i64, dp, i64, dp, i32]),
            "orc_node_gather": (None, [ip, i64, i64, ip, dp, dp, d, dp]),
            "orc_mobility_bcc": (None, [dp, d, ip, i64, i64, ip, dp, dp, dp, ip, i64, dp, dp,
                                        dp]),
            "orc_u_tilde": (None, [dp, i64, ip, dp, dp, C.POINTER(C.c_int32), i64, dp, dp, dp,
                                   dp, i64, dp]),
            "orc_minimum_distance": (None, [dp] * 9),
            "orc_detect_collision": (None, [d, i64, i64, ip, ip, i64, dp, dp, dp,
                                            C.POINTER(C.c_int32), ip, i64, ip, dp]),
            "orc_set_libm": (None, [i32]),
            "orc_segseg_near": (None, [dp] + net + [ip, i32, i32, dp, i32]),
            "orc_segseg_near_once": (None, [dp] + net + [ip, i32, i32, ip, dp]),
            "orc_segseg_mixed": (i64, [i32, i32, d, d, dp, dp] + net + [dp, i32]),
            "orc_segseg_rows_grouped": (None, [dp] + net + [ip, i64, ip, i64, dp, i32]),
            "orc_segseg_rows_binned": (None, [dp] + net + [ip, i64, dp, i32, dp, i32]),
            "orc_segseg_par_rows": (i64, [dp] + net + [ip, i64, dp, i32]),
            "orc_segseg_rows_triangle": (None, [dp] + net + [ip, i64, dp, i32]),
            "orc_segseg_tuned_rows": (i64, [dp] + net + [ip, i64, dp, i32]),
            "orc_libm_log": (d, [d]),
            "orc_libm_atan": (d, [d]),
            "orc_libm_vec": (None, [i32, dp, i64, dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(_lib, name)
            f.restype = res
            f.argtypes = args
    return _lib


def _d(a):
    return a.ctypes.data_as(dp)


def _i(a):
    return a.ctypes.data_as(ip)


def _u(a):
    return a.ctypes.data_as(up)


def _f(a, dt=np.float64):
    return np.asfortranarray(a, dtype=dt)


def set_libm(glibc: bool):
    """Use glibc's log/atan (True) or the fdlibm-method ones the kernels use."""
    lib().orc_set_libm(1 if glibc else 0)


def libm_log(x):
    return lib().orc_libm_log(float(x))


def libm_atan(x):
    return lib().orc_libm_atan(float(x))


def libm_vec(which, x):
    """log (which 0), atan (1), sqrt (2) of every x, or x[0::2] / x[1::2] (3)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.size // 2 if which == 3 else x.size
    y = np.empty(n)
    lib().orc_libm_vec(which, _d(x), n, _d(y))
    return y


def seg_consts(dln, mat):
    """(aSq, mu4pi, mu8pi, aSq mu8pi, mu4pinu, aSq mu4pinu): DislocationForces.jl:670-718."""
    aSq = dln.coreRadSq
    return np.array([aSq, mat.μ4π, mat.μ8π, aSq * mat.μ8π, mat.μ4πν, aSq * mat.μ4πν])


def self_consts(dln, mat):
    return np.array([dln.coreRad, dln.coreRadSq, mat.μ4π, mat.νomνInv, dln.coreEnergy, mat.omνInv])


def sigma_consts(dln, mat):
    return np.array([dln.coreRadSq, mat.μ4π, mat.μ4πν, mat.νμ4πν, dln.coreRadSq * mat.μ8π])


class _Net:
    def __init__(self, net):
        self.numSeg = int(net.numSeg[0])
        self.segIdx = _f(net.segIdx, np.int64)
        self.ld = self.segIdx.shape[0]
        self.coord = _f(net.coord)
        self.bVec = _f(net.bVec)
        self.ext = np.ascontiguousarray(net.extSeg, dtype=np.uint8)


def _idx(idx):
    if idx is None:
        return None, 0
    a = np.atleast_1d(np.asarray(list(idx) if isinstance(idx, range) else idx, dtype=np.int64))
    return np.ascontiguousarray(a), len(a)


def pair(b1, n11, n12, b2, n21, n22, consts, prec=0):
    """F1..F4 (4 x 3) and whether the near-parallel branch ran."""
    out = np.zeros(12)
    v = [np.ascontiguousarray(x, dtype=np.float64) for x in (b1, n11, n12, b2, n21, n22)]
    k = np.ascontiguousarray(consts, dtype=np.float64)
    par = lib().orc_segseg_pair(prec, _d(k), *[_d(x) for x in v], _d(out))
    return out.reshape(4, 3), bool(par)


def count_pair(b1, n11, n12, b2, n21, n22, consts):
    out = np.zeros(7)
    v = [np.ascontiguousarray(x, dtype=np.float64) for x in (b1, n11, n12, b2, n21, n22)]
    k = np.ascontiguousarray(consts, dtype=np.float64)
    par = lib().orc_count_segseg_pair(_d(k), *[_d(x) for x in v], _d(out))
    keys = ("flops", "add", "mul", "div", "sqrt", "log", "atan")
    return dict(zip(keys, out.tolist())), bool(par)


def count_sigma_tilde(sc, b, n1, n2, x):
    out = np.zeros(7)
    v = [np.ascontiguousarray(a, dtype=np.float64) for a in (sc, b, n1, n2, x)]
    lib().orc_count_sigma_tilde(*[_d(a) for a in v], _d(out))
    return dict(zip(("flops", "add", "mul", "div", "sqrt", "log", "atan"), out.tolist()))


def count_self_force(sc, tv, b):
    out = np.zeros(7)
    v = [np.ascontiguousarray(a, dtype=np.float64) for a in (sc, tv, b)]
    lib().orc_count_self_force(*[_d(a) for a in v], _d(out))
    return dict(zip(("flops", "add", "mul", "div", "sqrt", "log", "atan"), out.tolist()))


def calcSegSegForce(dln, mat, net, idx=None, prec=0):
    """Out-of-place seg-seg force: 3 x 2 x numSeg (serial, :778-829) or
    3 x 2 x length(idx) (rows only, :831-877)."""
    n = _Net(net)
    K = seg_consts(dln, mat)
    if idx is None:
        out = np.zeros((3, 2, n.numSeg), order="F")
        lib().orc_segseg_ser(prec, _d(K), n.numSeg, _i(n.segIdx), n.ld, _d(n.coord), _d(n.bVec),
                             _u(n.ext), _d(out))
        return out
    ix, ni = _idx(idx)
    out = np.zeros((3, 2, ni), order="F")
    lib().orc_segseg_idx(_d(K), n.numSeg, _i(n.segIdx), n.ld, _d(n.coord), _d(n.bVec),
                         _u(n.ext), _i(ix), ni, _d(out), 1)
    return out


def calcSegSegForcePar(dln, mat, net, nthreads):
    """The reference's threaded design (:719-777): static chunks + atomic adds."""
    n = _Net(net)
    K = seg_consts(dln, mat)
    out = np.zeros((3, 2, n.numSeg), order="F")
    lib().orc_segseg_par_faithful(_d(K), n.numSeg, _i(n.segIdx), n.ld, _d(n.coord), _d(n.bVec),
                                  _u(n.ext), _d(out), int(nthreads))
    return out


def segseg_rows_binned(dln, mat, net, rows, edges, nthreads=None):
    """Row forces split by source midpoint distance: 6 x nbins x len(rows)
    (F1 then F2 of each row segment, from sources in each distance bin)."""
    n = _Net(net)
    K = seg_consts(dln, mat)
    r = np.ascontiguousarray(rows, dtype=np.int64)
    e = np.ascontiguousarray(edges, dtype=np.float64)
    nb = len(e) - 1
    out = np.zeros((6, nb, len(r)), order="F")
    lib().orc_segseg_rows_binned(_d(K), n.numSeg, _i(n.segIdx), n.ld, _d(n.coord), _d(n.bVec),
                                 _u(n.ext), _i(r), len(r), _d(e), nb, _d(out),
                                 int(nthreads or os.cpu_count() or 1))
    return out


def segseg_rows_grouped(dln, mat, net, rows, group, ngroups, nthreads=None):
    """Row forces split by source group (group[j] for source j, -1 to skip):
    6 x ngroups x len(rows)."""
    n = _Net(net)
    K = seg_consts(dln, mat)
    r = np.ascontiguousarray(rows, dtype=np.int64)
    gr = np.ascontiguousarray(group, dtype=np.int64)
    out = np.zeros((6, ngroups, len(r)), order="F")
    lib().orc_segseg_rows_grouped(_d(K), n.numSeg, _i(n.segIdx), n.ld, _d(n.coord), _d(n.bVec),
                                  _u(n.ext), _i(r), len(r), _i(gr), ngroups, _d(out),
                                  int(nthreads or os.cpu_count() or 1))
    return out


def calcSegSegForceMixed(dln, mat, net, prec=2, nq=3, farDist=None, sinMin=1e-3,
                         nthreads=None):
    """Mixed rule: far (midpoint distance > farDist, default 5 a, and
    |t1 x t2| > sinMin) pairs by quadrature in float (prec 2) or double (0);
    others FP64 analytic.  Returns (3 x 2 x numSeg, pairs on FP64)."""
    n = _Net(net)
    K = seg_consts(dln, mat)
    sc = sigma_consts(dln, mat)
    fd = 5 * dln.coreRad if farDist is None else farDist
    out = np.zeros((3, 2, n.numSeg), order="F")
    ex = lib().orc_segseg_mixed(int(prec), int(nq), float(fd), float(sinMin), _d(K), _d(sc),
                                n.numSeg, _i(n.segIdx), n.ld, _d(n.coord), _d(n.bVec), _u(n.ext),
                                _d(out), int(nthreads or os.cpu_count() or 1))
    return out, int(ex)


def segseg_near(dln, mat, net, cellOf, G, rn, nthreads=None):
    """Near-field rows of the hybrid split (pairs whose cells are within
    Chebyshev distance rn, row segment first): 3 x 2 x numSeg."""
    n = _Net(net)
    K = seg_consts(dln, mat)
    co = np.ascontiguousarray(cellOf, dtype=np.int64)
    out = np.zeros((3, 2, n.numSeg), order="F")
    lib().orc_segseg_near(_d(K), n.numSeg, _i(n.segIdx), n.ld, _d(n.coord), _d(n.bVec), _u(n.ext),
                          _i(co), int(G), int(rn), _d(out), int(nthreads or os.cpu_count() or 1))
    return out


def segseg_near_once(dln, mat, net, cellOf, G, rn, rank):
    """The near pairs of segseg_near, each evaluated once with the segment of
    lower rank first, forces added to both: 3 x 2 x numSeg."""
    n = _Net(net)
    K = seg_consts(dln, mat)
    co = np.ascontiguousarray(cellOf, dtype=np.int64)
    rk = np.ascontiguousarray(rank, dtype=np.int64)
    out = np.zeros((3, 2, n.numSeg), order="F")
    lib().orc_segseg_near_once(_d(K), n.numSeg, _i(n.segIdx), n.ld, _d(n.coord), _d(n.bVec),
                               _u(n.ext), _i(co), int(G), int(rn), _i(rk), _d(out))
    return out


def segseg_par_rows(dln, mat, net, rows, nthreads):
    """The reference's threaded design on a sample of outer rows (each row i
    against all j > i); returns (segForce 3 x 2 x numSeg, pairs evaluated)."""
    n = _Net(net)
    K = seg_consts(dln, mat)
    r = np.ascontiguousarray(rows, dtype=np.int64)
    out = np.zeros((3, 2, n.numSeg), order="F")
    pairs = lib().orc_segseg_par_rows(_d(K), n.numSeg, _i(n.segIdx), n.ld, _d(n.coord),
                                      _d(n.bVec), _u(n.ext), _i(r), len(r), _d(out),
                                      int(nthreads))
    return out, int(pairs)


def segseg_rows_triangle(dln, mat, net, rows, nthreads=16):
    """Seg-seg forces of the all-pairs triangle (calcSegSegForceSer!'s
    orientation) on the sampled 1-based rows only: 3 x 2 x len(rows)."""
    n = _Net(net)
    K = seg_consts(dln, mat)
    r = np.ascontiguousarray(rows, dtype=np.int64)
    out = np.zeros((3, 2, len(r)), order="F")
    lib().orc_segseg_rows_triangle(_d(K), n.numSeg, _i(n.segIdx), n.ld, _d(n.coord), _d(n.bVec),
                                   _u(n.ext), _i(r), len(r), _d(out), int(nthreads))
    return out


def calcSelfForce(dln, mat, net, idx=None):
    n = _Net(net)
    ix, ni = _idx(idx)
    cnt = ni if ix is not None else n.numSeg
    out = np.zeros((3, 2, cnt), order="F")
    lib().orc_self_force(_d(self_consts(dln, mat)), n.numSeg, _i(n.segIdx), n.ld, _d(n.coord),
                         _d(n.bVec), _u(n.ext), _i(ix) if ix is not None else None, ni, _d(out), 1)
    return out[:, 0, :].copy(order="F"), out[:, 1, :].copy(order="F")


def _mesh_args(mesh, fd):
    return (_d(_f(mesh.C)), _d(_f(mesh.coord)), _i(_f(mesh.connectivity, np.int64)),
            int(mesh.mx), int(mesh.my), int(mesh.mz), C.c_double(mesh.w), C.c_double(mesh.h),
            C.c_double(mesh.d), _d(np.ascontiguousarray(fd.uHat, dtype=np.float64)))


def calc_σHat(mesh, fd, x0):
    out = np.zeros(9)
    x = np.ascontiguousarray(x0, dtype=np.float64)
    keep = (_f(mesh.C), _f(mesh.coord), _f(mesh.connectivity, np.int64),
            np.ascontiguousarray(fd.uHat, dtype=np.float64))
    lib().orc_sigma_hat(_d(keep[0]), _d(keep[1]), _i(keep[2]), int(mesh.mx), int(mesh.my),
                        int(mesh.mz), C.c_double(mesh.w), C.c_double(mesh.h), C.c_double(mesh.d), _d(keep[3]),
                        _d(x), _d(out))
    return out.reshape(3, 3, order="F")


def calcPKForce(mesh, fd, net, idx=None):
    n = _Net(net)
    ix, ni = _idx(idx)
    cnt = ni if ix is not None else n.numSeg
    out = np.zeros((3, cnt), order="F")
    keep = (_f(mesh.C), _f(mesh.coord), _f(mesh.connectivity, np.int64),
            np.ascontiguousarray(fd.uHat, dtype=np.float64))
    lib().orc_pk_force(_d(keep[0]), _d(keep[1]), _i(keep[2]), int(mesh.mx), int(mesh.my),
                       int(mesh.mz), C.c_double(mesh.w), C.c_double(mesh.h), C.c_double(mesh.d), _d(keep[3]),
                       n.numSeg, _i(n.segIdx), n.ld, _d(n.coord), _d(n.bVec), _u(n.ext),
                       _i(ix) if ix is not None else None, ni, _d(out), 1, 0)
    return out


def calcSegForce(dln, mat, mesh, fd, net, idx=None):
    """segseg + self + PK/2 (DislocationForces.jl:23-46)."""
    ss = calcSegSegForce(dln, mat, net, idx)
    s1, s2 = calcSelfForce(dln, mat, net, idx)
    pk = calcPKForce(mesh, fd, net, idx)
    ss[:, 0, :] += s1 + 0.5 * pk
    ss[:, 1, :] += s2 + 0.5 * pk
    return ss


def calc_σTilde(x0, dln, mat, net, idx=None, nthreads=None):
    """6 x P stress (xx, yy, zz, xy, xz, yz): DislocationFEM.jl:24-41,86-303."""
    n = _Net(net)
    x = _f(np.asarray(x0, dtype=np.float64).reshape(3, -1, order="F"))
    P = x.shape[1]
    ix, ni = _idx(idx)
    out = np.zeros((6, P), order="F")
    nt = nthreads or os.cpu_count() or 1
    lib().orc_sigma_tilde(_d(sigma_consts(dln, mat)), n.numSeg, _i(n.segIdx), n.ld, _d(n.coord),
                          _d(n.bVec), _i(ix) if ix is not None else None, ni, _d(x), P, _d(out),
                          int(nt))
    return out


def nodeForce(mat, net, segForce=None):
    """Node-force gather of dlnMobility! (DislocationMobility.jl:220-255)."""
    sf = _f(net.segForce if segForce is None else segForce)
    conn = _f(net.connectivity, np.int64)
    links = _f(net.links, np.int64)
    coord = _f(net.coord)
    numNode = int(net.numNode[0])
    out = np.zeros((3, numNode), order="F")
    lib().orc_node_gather(_i(conn), conn.shape[0], numNode, _i(links), _d(coord), _d(sf),
                          C.c_double(mat.σPN), _d(out))
    return out


def drag_coeffs(dln):
    """(edge, screw, climb, line) of dlnParams.dragCoeffs
    (DislocationMobility.jl:25-28)."""
    d = dln.dragCoeffs
    return np.array([d["edge"], d["screw"], d["climb"], d["line"]], dtype=np.float64)


def dlnMobility(dln, mat, net, idx=None, segForce=None, with_drag=False):
    """dlnMobility for mobBCC (DislocationMobility.jl:16-171): (nodeForce,
    nodeVel), 3 x n with column k for the k-th node of idx (all nodes when
    idx is None); with_drag also returns the 3 x 3 x n drag tensors."""
    sf = _f(net.segForce if segForce is None else segForce)
    conn = _f(net.connectivity, np.int64)
    links = _f(net.links, np.int64)
    coord = _f(net.coord)
    bVec = _f(net.bVec)
    numNode = int(net.numNode[0])
    ix = None if idx is None else np.ascontiguousarray(np.atleast_1d(idx), dtype=np.int64)
    n = numNode if ix is None else len(ix)
    F = np.zeros((3, n), order="F")
    V = np.zeros((3, n), order="F")
    D = np.zeros((3, 3, n), order="F") if with_drag else None
    lib().orc_mobility_bcc(_d(drag_coeffs(dln)), C.c_double(mat.σPN), _i(conn), conn.shape[0],
                           numNode, _i(links), _d(coord), _d(bVec), _d(sf),
                           _i(ix) if ix is not None else None, n, _d(F), _d(V),
                           _d(D) if with_drag else None)
    return (F, V, D) if with_drag else (F, V)


FIXED_LABELS = (2, 4, 5)   # intFixDln, srfFixDln, extDln (DislocationType.jl:24-32)


def deriv_(dln, mat, mesh, fd, net):
    """deriv! (IntegrationProcessing.jl:13-62): calcSegForce! (no PK when
    mesh is None), dlnMobility! and zero velocity of fixed and external
    nodes; fills net.segForce, nodeForce, nodeVel.  Surface-mobile nodes are
    refused: the reference's correction for them (:34-60) indexes label with
    a tuple, which Julia rejects."""
    nn, ns = int(net.numNode[0]), int(net.numSeg[0])
    lab = np.asarray(net.label[:nn])
    if np.any(lab == 3):
        raise ValueError("srfMobDln nodes: the reference's surface correction cannot run")
    # the serial triangle: a fixed summation order, so that repeated runs
    # take the same adaptive steps (the threaded one adds atomically)
    sf = calcSegSegForce(dln, mat, net)
    s1, s2 = calcSelfForce(dln, mat, net)
    if mesh is not None:
        pk = calcPKForce(mesh, fd, net)
        s1 = s1 + 0.5 * pk
        s2 = s2 + 0.5 * pk
    sf[:, 0, :] += s1
    sf[:, 1, :] += s2
    net.segForce[:, :, :ns] = sf
    F, V = dlnMobility(dln, mat, net)
    V[:, np.isin(lab, FIXED_LABELS)] = 0.0
    net.nodeForce[:, :nn] = F
    net.nodeVel[:, :nn] = V


def integrate_(ip, itime, dln, mat, mesh, fd, net):
    """integrate! for AdaptiveEulerTrapezoid (IntegrationProcessing.jl:81-168):
    moves net.coord and returns (dt, time, step) as the reference does (the
    returned dt is the next step's proposal, and time advances by it)."""
    import math
    nn = int(net.numNode[0])
    if nn == 0:
        return itime.dt, itime.time, itime.step
    deriv_(dln, mat, mesh, fd, net)
    c0 = net.coord[:, :nn].copy()
    v0 = net.nodeVel[:, :nn].copy()
    dt = max(itime.dt, ip.dtmin)
    dtOld, dtOldGood, convergent, counter = 0.0, False, False, 1
    while not convergent:
        net.coord[:, :nn] = c0 + v0 * dt
        deriv_(dln, mat, mesh, fd, net)
        distance = net.coord[:, :nn] - c0
        maxDist = np.max(np.abs(distance))
        err = distance - (net.nodeVel[:, :nn] + v0) / 2 * dt
        maxErr = np.max(np.abs(err))
        if dt <= ip.dtmin:
            counter = ip.maxiter + 1
        elif maxDist < ip.abstol and maxErr < ip.reltol:
            dtOld = dt
            factor = ip.maxchange * math.pow(
                1 / (1 + (math.pow(ip.maxchange, ip.exponent) - 1) * (maxErr / ip.reltol)),
                1 / ip.exponent)
            dtOldGood = True
            dt = min(dt * factor, ip.dtmax)
            counter += 1
        elif dtOldGood:
            dt = dtOld
            counter = ip.maxiter
        else:
            dt = dt / 2
        if counter > ip.maxiter or dt == ip.dtmax:
            convergent = True
    return dt, itime.time + dt, itime.step + 1


def u_tilde_points(mat, net, mesh, points):
    """Displacement of the network at the points (3 x P): the sum of
    _calc_uTilde! (DislocationFEM.jl:381-600) before it is scattered into
    dofs; column i belongs to point i."""
    x = _f(np.asarray(points, dtype=np.float64).reshape(3, -1, order="F"))
    P = x.shape[1]
    out = np.zeros((3, P), order="F")
    links = _f(net.links, np.int64)
    coord = _f(net.coord)
    bVec = _f(net.bVec)
    lab = np.ascontiguousarray(net.label, dtype=np.int32)
    k2 = np.array([mat.omνInv8π, mat.om2νomνInv8π])
    lib().orc_u_tilde(_d(k2), int(net.numSeg[0]), _i(links), _d(coord), _d(bVec),
                      lab.ctypes.data_as(C.POINTER(C.c_int32)), coord.shape[1],
                      _d(_f(mesh.vertices)), _d(_f(mesh.faceNorm)), _d(_f(mesh.faceMidPt)), _d(x),
                      P, _d(out))
    return out


def calc_uTilde(mesh, boundaries, mat, net):
    """calc_uTilde (DislocationFEM.jl:369-379): the displacements at the
    nodes uGammaDln, interleaved (x, y, z of each node in turn)."""
    pts = mesh.coord[:, np.asarray(boundaries.uGammaDln) - 1]
    return u_tilde_points(mat, net, mesh, pts).ravel(order="F")


def calc_uTilde_(fd, mesh, boundaries, mat, net):
    """calc_uTilde! (DislocationFEM.jl:355-367): zero fd.uTilde, then add
    value k of the interleaved displacements into dof uDofsDln[k] (the
    reference's mapping, whose dof list is ordered x, then y, then z)."""
    v = calc_uTilde(mesh, boundaries, mat, net)
    fd.uTilde[:] = 0.0
    np.add.at(fd.uTilde, np.asarray(boundaries.uDofsDln) - 1, v)
    return fd.uTilde


COLLISION_TYPES = {0: "null", 1: "hinge", 2: "twoLine"}


def minimumDistance(x0, x1, y0, y1, vx0, vx1, vy0, vy1):
    """(distSq, dDistSqDt, L1, L2) of two moving segments (Misc.jl:314-367)."""
    v = [np.ascontiguousarray(a, dtype=np.float64) for a in (x0, x1, y0, y1, vx0, vx1, vy0, vy1)]
    out = np.zeros(4)
    lib().orc_minimum_distance(*[_d(a) for a in v], _d(out))
    return tuple(float(x) for x in out)


def detectCollision(dln, net, skipSegs=()):
    """detectCollision (DislocationTopologyCollision.jl:1-222): (collision,
    type, n1s1, n2s1, n1s2, n2s2, s1, s2, L1, L2)."""
    links = _f(net.links, np.int64)
    conn = _f(net.connectivity, np.int64)
    coord = _f(net.coord)
    vel = _f(net.nodeVel)
    bVec = _f(net.bVec)
    lab = np.ascontiguousarray(net.label, dtype=np.int32)
    sk = np.ascontiguousarray(np.asarray(list(skipSegs), dtype=np.int64).reshape(-1, 2))
    res = np.zeros(8, dtype=np.int64)
    L = np.zeros(2)
    lib().orc_detect_collision(C.c_double(dln.collisionDistSq), int(net.numNode[0]),
                               int(net.numSeg[0]), _i(links), _i(conn), conn.shape[0], _d(coord),
                               _d(vel), _d(bVec), lab.ctypes.data_as(C.POINTER(C.c_int32)),
                               _i(sk), len(sk), _i(res), _d(L))
    return (bool(res[0]), COLLISION_TYPES[int(res[1])], *[int(x) for x in res[2:]],
            float(L[0]), float(L[1]))


def segseg_tuned_rows(dln, mat, net, rows, nthreads):
    """A tuned CPU design on a sample of outer rows (dynamic schedule,
    per-thread accumulators); returns (segForce 3 x 2 x numSeg, pairs)."""
    n = _Net(net)
    K = seg_consts(dln, mat)
    r = np.ascontiguousarray(rows, dtype=np.int64)
    out = np.zeros((3, 2, n.numSeg), order="F")
    pairs = lib().orc_segseg_tuned_rows(_d(K), n.numSeg, _i(n.segIdx), n.ld, _d(n.coord),
                                        _d(n.bVec), _u(n.ext), _i(r), len(r), _d(out),
                                        int(nthreads))
    return out, int(pairs)
