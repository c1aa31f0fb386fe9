"""The allocation-free calc_σTilde entry point (ddd_sigma_tilde_noalloc)
against the allocating one (ddd_sigma_tilde, DislocationFEM.jl:24-303)
through the C-ABI: inputs and outputs in caller-owned device memory, the
library's temporaries in its fixed per-GPU scratch.  Two calls in a row, the
second on a smaller input, each bit-identical to the allocating entry point
(same chunking, same sums) and within 1e-10 of it normwise; accumulation;
an input beyond the scratch is refused before any work."""
import ctypes as C

import numpy as np
import pytest

from dddhip import _lib as L
from dddhip import forces as G
from dddhip.generate import random_loop_network, face_grid_points
from dddhip.types import DislocationParameters, MaterialParameters
from metrics import normwise

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


class _Dev:
    """Caller-owned device memory from the HIP runtime the library uses
    (hipMalloc / hipMemcpy through ctypes), in column-major (Julia) order."""

    def __init__(self, a, dtype):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.host = np.ascontiguousarray(np.asarray(a, dtype=dtype).ravel(order="F"))
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), C.c_size_t(max(self.host.nbytes, 8))) == 0
        if self.host.nbytes:
            assert self.hip.hipMemcpy(self.p, self.host.ctypes.data_as(C.c_void_p),
                                      C.c_size_t(self.host.nbytes), 1) == 0  # host to device

    def get(self):
        assert self.hip.hipDeviceSynchronize() == 0
        out = np.empty_like(self.host)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p,
                                  C.c_size_t(out.nbytes), 2) == 0  # device to host
        return out

    def __del__(self):
        if self.p:
            self.hip.hipFree(self.p)


def _noalloc(ctx, dln, mat, net, x, idx=None, sigma=None, flags=0):
    k = L.consts(dln, mat)
    segIdx = _Dev(np.asfortranarray(net.segIdx), np.int64)
    coord = _Dev(np.asfortranarray(net.coord), np.float64)
    bVec = _Dev(np.asfortranarray(net.bVec), np.float64)
    x0 = _Dev(np.asfortranarray(x), np.float64)
    P = x.shape[1]
    out = _Dev(np.zeros((6, P)) if sigma is None else np.asfortranarray(sigma), np.float64)
    ix = None if idx is None else _Dev(np.asarray(idx, dtype=np.int64), np.int64)
    rc = L.lib().ddd_sigma_tilde_noalloc(
        ctx.h, C.byref(k), int(net.numSeg[0]), segIdx.p, net.segIdx.shape[0], coord.p,
        net.coord.shape[1], bVec.p, net.bVec.shape[1], None if ix is None else ix.p,
        0 if ix is None else len(idx), x0.p, P, out.p, flags)
    ctx.check(rc)
    return out.get().reshape(6, P, order="F")


def test_matches_allocating_entry_point_twice():
    dln = DislocationParameters(coreRad=90.78, coreEnergy=1.0)
    mat = MaterialParameters(μ=1.0, ν=0.28)
    ctx = L.default_context()
    big = random_loop_network(64, seed=0xB7)
    xb = face_grid_points(32)
    _log(f"noalloc call 1: {int(big.numSeg[0])} segments x {xb.shape[1]} points")
    a = _noalloc(ctx, dln, mat, big, xb)
    ref = G.calc_σTilde(xb, dln, mat, big, ctx=ctx)
    assert np.all(np.isfinite(a))
    assert np.array_equal(a, ref) and normwise(a, ref) <= 1e-10
    # the second call is smaller: fewer segments (an idx subset of another
    # network) and fewer points; the scratch keeps the first call's contents
    small = random_loop_network(8, seed=0xB8)
    xs = face_grid_points(8)[:, ::3]
    idx = np.arange(1, int(small.numSeg[0]) + 1, 2)
    _log(f"noalloc call 2: {len(idx)} segments x {xs.shape[1]} points")
    b = _noalloc(ctx, dln, mat, small, xs, idx=idx)
    refb = G.calc_σTilde(xs, dln, mat, small, idx, ctx=ctx)
    assert np.array_equal(b, refb) and normwise(b, refb) <= 1e-10
    # accumulate into a given sigma
    s0 = np.random.default_rng(3).standard_normal((6, xs.shape[1]))
    c = _noalloc(ctx, dln, mat, small, xs, idx=idx, sigma=s0, flags=L.DDD_ACCUMULATE)
    assert normwise(c, s0 + refb) <= 1e-12


def test_refuses_beyond_scratch_and_empty():
    dln = DislocationParameters(coreRad=90.78, coreEnergy=1.0)
    mat = MaterialParameters(μ=1.0, ν=0.28)
    ctx = L.default_context()
    net = random_loop_network(2, seed=0xB9)
    k = L.consts(dln, mat)
    # 1.5 M points need 9 M doubles of partials: refused before any work
    P = 3 * (1 << 19)
    segIdx = _Dev(np.asfortranarray(net.segIdx), np.int64)
    coord = _Dev(np.asfortranarray(net.coord), np.float64)
    bVec = _Dev(np.asfortranarray(net.bVec), np.float64)
    x0 = _Dev(np.zeros(3 * P), np.float64)
    sig = _Dev(np.zeros(6 * P), np.float64)
    rc = L.lib().ddd_sigma_tilde_noalloc(ctx.h, C.byref(k), int(net.numSeg[0]), segIdx.p,
                                         net.segIdx.shape[0], coord.p, net.coord.shape[1],
                                         bVec.p, net.bVec.shape[1], None, 0, x0.p, P, sig.p, 0)
    assert rc == L.DDD_ERR_ARG and b"scratch" in L.lib().ddd_last_error(ctx.h)
    assert np.array_equal(sig.get(), np.zeros(6 * P))
    # no segments: sigma = 0 (or unchanged when accumulating)
    x = face_grid_points(4)
    e = random_loop_network(1, seed=0xBA)
    e.numSeg[0] = 0
    z = _noalloc(ctx, dln, mat, e, x, sigma=np.ones((6, x.shape[1])))
    assert np.array_equal(z, np.zeros_like(z))
    o = _noalloc(ctx, dln, mat, e, x, sigma=np.ones((6, x.shape[1])), flags=L.DDD_ACCUMULATE)
    assert np.array_equal(o, np.ones_like(o))
