"""GPU: the all-pairs evaluation in chunks of bounded memory.

The triangle of 64 x 64 pair tiles is evaluated chunk by chunk (whole
column blocks whose partial sums and near-parallel lists fit the context's
budget, default 32 GiB), each chunk reduced into segForce in a fixed order.
Small budgets force many chunks on networks the oracle checks
(calcSegSegForcePar, DislocationForces.jl:719-777, 1e-10 per node); a
524,288-segment network runs within the default budget and is checked by
size-independent properties: 48 sampled rows against the oracle's
triangle in the same orientation (:794-825) and the balance of forces over
closed loops."""
import numpy as np
import pytest

import pyoracle as O
from dddhip import forces as G
from dddhip.generate import clustered_loop_network, random_loop_network
from dddhip.types import DislocationParameters, MaterialParameters
from metrics import normwise, per_node_rel

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


def _prm():
    return (DislocationParameters(coreRad=90.78, coreEnergy=1.0),
            MaterialParameters(μ=1.0, ν=0.28))


@pytest.fixture
def small_chunks():
    # 1 MiB: one or a few super-columns of 4 x 4-tile super-tiles (24 KiB of
    # partial sums each) and their near-parallel pairs
    G.setChunkBytes(1 << 20)
    yield
    G.setChunkBytes(0)


def test_many_chunks_match_oracle_and_one_chunk(small_chunks):
    dln, mat = _prm()
    net = random_loop_network(512, seed=0xB0)
    _log("chunked FP64: 16,384 segments, 1 MiB chunks")
    g = G.calcSegSegForce(dln, mat, net)
    res = G.ResidentNetwork(dln, mat, net)
    res.seg_force()
    nbytes, chunks = res.footprint()
    _log(f"chunks {chunks}, device bytes {nbytes}")
    assert chunks >= 16  # of the 64 super-columns
    o = O.calcSegSegForcePar(dln, mat, net, 16)
    assert per_node_rel(g, o)[0] <= 1e-10
    assert np.array_equal(g, G.calcSegSegForce(dln, mat, net))  # repeatable
    # the mixed mode shares the chunks (its pairs are evaluated identically)
    m, fp64 = G.calcSegSegForceMixed(dln, mat, net)
    G.setChunkBytes(0)
    one = G.calcSegSegForce(dln, mat, net)
    m1, fp64_1 = G.calcSegSegForceMixed(dln, mat, net)
    assert normwise(g, one) <= 1e-13  # summation order only
    assert normwise(m, m1) <= 1e-13 and fp64 == fp64_1 > 0


def test_many_chunks_near_parallel_pileup(small_chunks):
    """Pile-ups put many near-parallel pairs in every chunk's list."""
    dln, mat = _prm()
    net = clustered_loop_network(numLoops=96, seed=0xE7, clusters=2, cluster_frac=0.9, sigma=120.0)
    _log(f"chunked pile-up: {int(net.numSeg[0])} segments")
    G.setChunkBytes(1 << 20)
    g = G.calcSegSegForce(dln, mat, net)
    o = O.calcSegSegForcePar(dln, mat, net, 16)
    assert per_node_rel(g, o)[0] <= 1e-10
    assert np.array_equal(g, G.calcSegSegForce(dln, mat, net))


def test_524288_segments_within_the_budget():
    dln, mat = _prm()
    net = random_loop_network(16384, seed=0xF0)
    n = int(net.numSeg[0])
    _log(f"plain FP64 at {n} segments")
    res = G.ResidentNetwork(dln, mat, net)
    res.seg_force()
    nbytes, chunks = res.footprint()
    f = res.download()
    _log(f"chunks {chunks}, device bytes {nbytes / 2**30:.2f} GiB")
    # unchunked, the super-tiles' partial sums (2,048 x 2,049 / 2 x 24 KiB = 48 GiB)
    # and the near-parallel list (~0.9e9 pairs x 110 B) would take about 150 GiB;
    # a chunk's buffer stays within the 32 GiB default budget; the per-tile
    # list offsets (0.4 GiB), the segment arrays and what earlier calls left
    # in the context take the rest
    assert chunks >= 2 and nbytes <= 34 * 2**30
    assert np.all(np.isfinite(f))
    # sampled rows against the oracle's triangle in the same orientation
    # (pair (min, max) of every pair); the resident total holds the self
    # force too
    rows = np.random.default_rng(1).choice(n, 48, replace=False) + 1
    r = O.segseg_rows_triangle(dln, mat, net, rows)
    s1, s2 = O.calcSelfForce(dln, mat, net, rows)
    r[:, 0] += s1
    r[:, 1] += s2
    assert per_node_rel(f[:, :, rows - 1], r)[0] <= 1e-10
    # closed loops in an infinite medium: the seg-seg forces sum to zero
    a1, a2 = O.calcSelfForce(dln, mat, net)
    ss = f.copy()
    ss[:, 0] -= a1
    ss[:, 1] -= a2
    tot = ss.sum(axis=(1, 2))
    assert np.linalg.norm(tot) <= 1e-9 * np.abs(ss).sum()
