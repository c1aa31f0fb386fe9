"""GPU: the mixed-precision path (FP32 quadrature for well-separated pairs,
FP64 for the rest) against the oracle's restatement of the same rule and
against the plain FP64 oracle.

Bars: against the FP64 oracle, per-node error <= 2e-5 max and <= 2e-6 RMS
(the 2-point quadrature's own accuracy; the oracle measures 2e-6 / 3e-7 on
8,192 segments); against the oracle's FP32 restatement of the mixed rule,
<= 1e-5 (FP32 rounding differences only); FP64-path pair counts agree to
1e-4."""
import numpy as np
import pytest

import pyoracle as O
import refnets as R
from dddhip import forces as G
from dddhip.generate import random_loop_network
from metrics import per_node_rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def prm():
    return R.params()


@pytest.mark.parametrize("loops,seed", [(20, 0xA0), (256, 0xC0)])
def test_mixed_against_oracle(prm, loops, seed):
    dln, mat = prm
    net = random_loop_network(loops, seed=seed)
    n = int(net.numSeg[0])
    print(f"mixed {n} segments", flush=True)
    g, ex = G.calcSegSegForceMixed(dln, mat, net)
    om, oex = O.calcSegSegForceMixed(dln, mat, net, prec=2, nq=2)
    ref = O.calcSegSegForcePar(dln, mat, net, 16)
    s1, s2 = O.calcSelfForce(dln, mat, net)
    for a in (g, om, ref):
        a[:, 0, :] += s1
        a[:, 1, :] += s2
    mx, rms = per_node_rel(g, ref)
    print(f"vs FP64 oracle: max {mx:.2e} rms {rms:.2e}; FP64 share {ex / (n * (n - 1) / 2):.3%}")
    assert mx <= 2e-5 and rms <= 2e-6
    assert per_node_rel(g, om)[0] <= 1e-5
    assert abs(ex - oex) <= max(2, 1e-4 * oex)


def test_mixed_all_fp64_when_far_threshold_huge(prm):
    """With every pair classified near, the mixed path evaluates every pair in
    FP64.  Its kernels are built with FMA contraction (an approximate path),
    so it differs from the uncontracted oracle only within the pair formula's
    conditioning."""
    dln, mat = prm
    net = random_loop_network(20, seed=0xA0)
    n = int(net.numSeg[0])
    g, ex = G.calcSegSegForceMixed(dln, mat, net, farDist=1e9)
    assert ex == n * (n - 1) // 2
    mx, rms = per_node_rel(g, O.calcSegSegForce(dln, mat, net))
    print(f"all-FP64 mixed vs strict oracle: max {mx:.2e} rms {rms:.2e}", flush=True)
    assert mx <= 1e-6 and rms <= 1e-7  # measured 1.2e-7 and 1.3e-8 on this network


def test_mixed_pileups(prm):
    """Mixed precision on pile-ups (many near-parallel pairs, listed per tile)
    against the FP64 oracle and the oracle's mixed restatement."""
    from dddhip.generate import clustered_loop_network
    dln, mat = prm
    net = clustered_loop_network(numLoops=96, seed=0xE8, clusters=2, cluster_frac=0.9,
                                 sigma=120.0)
    g, ex = G.calcSegSegForceMixed(dln, mat, net)
    om, oex = O.calcSegSegForceMixed(dln, mat, net, prec=2, nq=2)
    ref = O.calcSegSegForcePar(dln, mat, net, 16)
    s1, s2 = O.calcSelfForce(dln, mat, net)
    for a in (g, om, ref):
        a[:, 0, :] += s1
        a[:, 1, :] += s2
    mx, rms = per_node_rel(g, ref)
    print(f"pile-ups vs FP64 oracle: max {mx:.2e} rms {rms:.2e}", flush=True)
    assert mx <= 2e-5 and rms <= 2e-6
    assert per_node_rel(g, om)[0] <= 1e-5
    assert abs(ex - oex) <= max(2, 1e-4 * oex)
