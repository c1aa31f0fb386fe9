"""ctypes binding of libdddhip.so (include/ddd_hip.h).

This is what a Julia ``ccall`` shim would bind; the structs mirror the C
declarations field for field.  There is no fallback: if the library or a
gfx950 device is missing, every call raises.
"""
from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parents[1]
LIB_PATH = PKG / "libdddhip.so"

DDD_OK, DDD_ERR_ARG, DDD_ERR_OOM, DDD_ERR_HIP, DDD_ERR_NO_DEVICE, DDD_ERR_NUMERIC, \
    DDD_ERR_INTERNAL = (0, -1, -2, -3, -4, -5, -6)
DDD_ACCUMULATE, DDD_OUT_BY_POSITION, DDD_PK_HALF_TO_NODES = 1, 2, 4

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int64)
up = C.POINTER(C.c_uint8)
i64 = C.c_int64


class ddd_consts(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "mu4pi", "mu8pi", "mu4pinu", "nuMu4pinu", "omNuInv", "nuOmNuInv", "coreRad",
        "coreRadSq", "coreEnergy", "sigmaPN")]


class ddd_drag(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("edge", "screw", "climb", "line")]


class ddd_integ(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("tmin", "tmax", "dtmin", "dtmax", "abstol", "reltol",
                                          "maxchange", "exponent")] + [("maxiter", C.c_int64)]


class ddd_itime(C.Structure):
    _fields_ = [("dt", C.c_double), ("time", C.c_double), ("step", C.c_int64)]


class ddd_collision(C.Structure):
    _fields_ = [("collision", C.c_int32), ("type", C.c_int32)] + [
        (n, C.c_int64) for n in ("n1s1", "n2s1", "n1s2", "n2s2", "s1", "s2")] + [
        ("L1", C.c_double), ("L2", C.c_double)]


class ddd_network(C.Structure):
    _fields_ = [("numNode", i64), ("numSeg", i64), ("segIdx", ip), ("segIdxLd", i64),
                ("coord", dp), ("coordCols", i64), ("bVec", dp), ("bVecCols", i64),
                ("extSeg", up), ("links", ip), ("linksCols", i64), ("connectivity", ip),
                ("connLd", i64), ("connCols", i64)]


class ddd_mesh(C.Structure):
    _fields_ = [("C", dp), ("coord", dp), ("numNode", i64), ("connectivity", ip),
                ("numElem", i64), ("mx", i64), ("my", i64), ("mz", i64), ("w", C.c_double),
                ("h", C.c_double), ("d", C.c_double), ("uHat", dp)]


DDD_MODE_FP64, DDD_MODE_MIXED, DDD_MODE_HYBRID = 0, 1, 2


class ddd_opts(C.Structure):
    _fields_ = [("mode", C.c_int), ("farDist", C.c_double), ("sinMin", C.c_double),
                ("farfield", C.c_void_p)]


class ddd_stats(C.Structure):
    _fields_ = [("pairs", C.c_uint64), ("pairsFP64", C.c_uint64), ("deviceBytes", C.c_uint64),
                ("chunks", C.c_uint64), ("tileMs", C.c_double)]


class ddd_train_opts(C.Structure):
    _fields_ = [("samples", i64), ("valSamples", i64), ("steps", C.c_int), ("batch", C.c_int),
                ("lr", C.c_float), ("lrFinal", C.c_float), ("seed", C.c_uint64),
                ("denFloor", C.c_float), ("nearFrac", C.c_float), ("nearShells", C.c_int),
                ("plainOffsets", C.c_int)]


class ddd_train_report(C.Structure):
    _fields_ = [("lossFirst", C.c_double), ("lossLast", C.c_double),
                ("relErrVal", C.c_double), ("relErrValNear", C.c_double),
                ("seconds", C.c_double), ("relErrShell", C.c_double * 8)]


class DDDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None
_lock = threading.Lock()


def lib():
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise DDDError(DDD_ERR_NO_DEVICE,
                               f"{LIB_PATH} is not built (run __graft_entry__.build())")
            L = C.CDLL(str(LIB_PATH))
            P = C.c_void_p
            sigs = {
                "ddd_version": (C.c_int, []),
                "ddd_ctx_create": (C.c_int, [C.POINTER(P), C.c_int]),
                "ddd_ctx_destroy": (C.c_int, [P]),
                "ddd_ctx_set_stream": (C.c_int, [P, P]),
                "ddd_ctx_stream": (P, [P]),
                "ddd_ctx_set_chunk_bytes": (C.c_int, [P, i64]),
                "ddd_last_error": (C.c_char_p, [P]),
                "ddd_segseg_force": (C.c_int, [P, C.POINTER(ddd_consts), C.POINTER(ddd_network),
                                               ip, i64, dp, i64, C.c_int]),
                "ddd_self_force": (C.c_int, [P, C.POINTER(ddd_consts), C.POINTER(ddd_network),
                                             ip, i64, dp, i64, C.c_int]),
                "ddd_pk_force": (C.c_int, [P, C.POINTER(ddd_network), C.POINTER(ddd_mesh), ip,
                                           i64, dp, i64, C.c_int]),
                "ddd_sigma_hat": (C.c_int, [P, C.POINTER(ddd_mesh), dp, i64, dp]),
                "ddd_selftest_ranged": (C.c_int, [P, C.c_int, dp, i64, dp]),
                "ddd_seg_force": (C.c_int, [P, C.POINTER(ddd_consts), C.POINTER(ddd_network),
                                            C.POINTER(ddd_mesh), ip, i64, dp, i64, C.c_int]),
                "ddd_sigma_tilde": (C.c_int, [P, C.POINTER(ddd_consts), C.POINTER(ddd_network),
                                              ip, i64, dp, i64, dp, C.c_int]),
                "ddd_sigma_tilde_noalloc": (C.c_int, [P, C.POINTER(ddd_consts), i64, P, i64, P,
                                                      i64, P, i64, P, i64, P, i64, P, C.c_int]),
                "ddd_node_force": (C.c_int, [P, C.POINTER(ddd_consts), C.POINTER(ddd_network),
                                             dp, i64, dp]),
                "ddd_net_upload_dynamics": (C.c_int, [P, C.POINTER(ddd_network),
                                                      C.POINTER(C.c_int32), C.POINTER(ddd_drag)]),
                "ddd_net_deriv": (C.c_int, [P, C.c_int]),
                "ddd_net_integrate": (C.c_int, [P, C.POINTER(ddd_integ), C.POINTER(ddd_itime),
                                                C.c_int]),
                "ddd_net_download_nodes": (C.c_int, [P, dp, dp, dp, i64]),
                "ddd_detect_collision": (C.c_int, [P, C.c_double, C.POINTER(ddd_network),
                                                   C.POINTER(C.c_int32), dp, ip, i64,
                                                   C.POINTER(ddd_collision)]),
                "ddd_u_tilde": (C.c_int, [P, C.c_double, C.c_double, C.POINTER(ddd_network),
                                          C.POINTER(C.c_int32), dp, dp, dp, dp, i64, dp,
                                          C.c_int]),
                "ddd_mobility_bcc": (C.c_int, [P, C.POINTER(ddd_consts), C.POINTER(ddd_drag),
                                               C.POINTER(ddd_network), dp, i64, ip, i64, dp, dp,
                                               i64, C.c_int]),
                "ddd_net_upload": (C.c_int, [P, C.POINTER(ddd_consts), C.POINTER(ddd_network),
                                             C.POINTER(ddd_mesh)]),
                "ddd_net_seg_force": (C.c_int, [P, C.c_int]),
                "ddd_net_stage": (C.c_int, [P, C.c_int, C.c_int]),
                "ddd_net_set_opts": (C.c_int, [P, C.POINTER(ddd_opts)]),
                "ddd_farfield_create": (C.c_int, [P, C.c_int, dp, C.c_double, C.c_int,
                                                  C.POINTER(P)]),
                "ddd_farfield_destroy": (C.c_int, [P, P]),
                "ddd_farfield_set_operator": (C.c_int, [P, P, C.POINTER(C.c_float)]),
                "ddd_farfield_debug": (C.c_int, [P, P, C.c_int, P, i64]),
                "ddd_net_stats": (C.c_int, [P, C.POINTER(ddd_stats)]),
                "ddd_segseg_force_opts": (C.c_int, [P, C.POINTER(ddd_consts),
                                                    C.POINTER(ddd_network), C.POINTER(ddd_opts),
                                                    dp, i64, C.c_int, C.POINTER(ddd_stats)]),
                "ddd_net_download": (C.c_int, [P, dp, i64]),
                "ddd_mlp_create": (C.c_int, [P, C.c_uint64, C.POINTER(P)]),
                "ddd_mlp_destroy": (C.c_int, [P, P]),
                "ddd_mlp_num_params": (i64, []),
                "ddd_mlp_params": (C.c_int, [P, P, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
                "ddd_mlp_loss_grad": (C.c_int, [P, P, C.c_int, C.POINTER(C.c_int),
                                                C.POINTER(C.c_float), C.POINTER(C.c_float),
                                                C.POINTER(C.c_float), C.c_float, dp,
                                                C.POINTER(C.c_float)]),
                "ddd_farfield_train": (C.c_int, [P, P, P, C.POINTER(ddd_consts),
                                                 C.POINTER(ddd_network),
                                                 C.POINTER(ddd_train_opts),
                                                 C.POINTER(ddd_train_report)]),
                "ddd_farfield_set_operator_mlp": (C.c_int, [P, P, P, C.POINTER(ddd_consts)]),
                "ddd_farfield_local_targets": (C.c_int, [P, P, C.POINTER(ddd_consts),
                                                         C.POINTER(ddd_network), C.c_int,
                                                         C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                         C.POINTER(C.c_float)]),
                "ddd_net_segforce_device": (P, [P]),
            }
            for name, (res, args) in sigs.items():
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
            _lib = L
    return _lib


class Context:
    """One device context (ddd_ctx_create / ddd_ctx_destroy)."""

    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        rc = lib().ddd_ctx_create(C.byref(self.h), int(device))
        if rc != DDD_OK:
            raise DDDError(rc, f"ddd_ctx_create(device={device}) failed: no usable gfx950 device")

    def check(self, rc):
        if rc != DDD_OK:
            raise DDDError(rc, lib().ddd_last_error(self.h).decode())

    def close(self):
        if self.h:
            lib().ddd_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = {}


def default_context(device: int = 0) -> Context:
    with _lock:
        ctx = _default.get(device)
    if ctx is None:
        ctx = Context(device)
        with _lock:
            _default[device] = ctx
    return ctx


def ptr(a, kind):
    if a is None:
        return None
    return a.ctypes.data_as({"d": dp, "i": ip, "u": up}[kind])


class NetView:
    """Holds Fortran-ordered views of a DislocationNetwork and the ddd_network
    struct that points into them (kept alive together)."""

    def __init__(self, net):
        self.segIdx = np.asfortranarray(net.segIdx, dtype=np.int64)
        self.coord = np.asfortranarray(net.coord, dtype=np.float64)
        self.bVec = np.asfortranarray(net.bVec, dtype=np.float64)
        self.ext = np.ascontiguousarray(net.extSeg, dtype=np.uint8)
        self.links = np.asfortranarray(net.links, dtype=np.int64)
        self.conn = np.asfortranarray(net.connectivity, dtype=np.int64)
        self.s = ddd_network(
            numNode=int(net.numNode[0]), numSeg=int(net.numSeg[0]), segIdx=ptr(self.segIdx, "i"),
            segIdxLd=self.segIdx.shape[0], coord=ptr(self.coord, "d"),
            coordCols=self.coord.shape[1], bVec=ptr(self.bVec, "d"), bVecCols=self.bVec.shape[1],
            extSeg=ptr(self.ext, "u"), links=ptr(self.links, "i"), linksCols=self.links.shape[1],
            connectivity=ptr(self.conn, "i"), connLd=self.conn.shape[0],
            connCols=self.conn.shape[1])


class MeshView:
    def __init__(self, mesh, fd):
        self.C = np.asfortranarray(mesh.C, dtype=np.float64)
        self.coord = np.asfortranarray(mesh.coord, dtype=np.float64)
        self.conn = np.asfortranarray(mesh.connectivity, dtype=np.int64)
        self.uHat = np.ascontiguousarray(fd.uHat, dtype=np.float64)
        self.s = ddd_mesh(C=ptr(self.C, "d"), coord=ptr(self.coord, "d"),
                          numNode=self.coord.shape[1], connectivity=ptr(self.conn, "i"),
                          numElem=self.conn.shape[1], mx=int(mesh.mx), my=int(mesh.my),
                          mz=int(mesh.mz), w=float(mesh.w), h=float(mesh.h), d=float(mesh.d),
                          uHat=ptr(self.uHat, "d"))


def consts(dlnParams, matParams) -> ddd_consts:
    return ddd_consts(mu4pi=matParams.μ4π, mu8pi=matParams.μ8π, mu4pinu=matParams.μ4πν,
                      nuMu4pinu=matParams.νμ4πν, omNuInv=matParams.omνInv,
                      nuOmNuInv=matParams.νomνInv, coreRad=dlnParams.coreRad,
                      coreRadSq=dlnParams.coreRadSq, coreEnergy=dlnParams.coreEnergy,
                      sigmaPN=matParams.σPN)


def drag(dlnParams):
    """ddd_drag from dlnParams.dragCoeffs; a missing entry fails as the
    reference's NamedTuple lookup does (DislocationMobility.jl:25-28)."""
    d = dlnParams.dragCoeffs
    return ddd_drag(float(d["edge"]), float(d["screw"]), float(d["climb"]), float(d["line"]))


def integ(intParams):
    if intParams.method != "AdaptiveEulerTrapezoid":
        raise NotImplementedError(f"integrator {intParams.method!r}")
    return ddd_integ(float(intParams.tmin), float(intParams.tmax), float(intParams.dtmin),
                     float(intParams.dtmax), float(intParams.abstol), float(intParams.reltol),
                     float(intParams.maxchange), float(intParams.exponent),
                     int(intParams.maxiter))
