"""GPU: the hybrid path's pieces against restatements.

* near field (all FP64, FMA-contracted, refined hardware reciprocals)
  against the oracle's near pairs in the same orientation to <= 1e-6 per
  node (the formula's conditioning);
  with the mixed rule, <= 2e-5;
* cell binning identical, moments to FP32 precision, the FFT correlation
  against a direct sum, and the far-field quadrature, with a random operator.
"""
import numpy as np
import pytest

import hybrid_ref as H
import pyoracle as O
import refnets as R
from dddhip import forces as G
from dddhip.generate import clustered_loop_network, random_loop_network
from dddhip.types import DislocationNetwork, intMobDln
from metrics import per_node_rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def prm():
    return R.params()


def _self(dln, mat, net):
    s1, s2 = O.calcSelfForce(dln, mat, net)
    out = np.zeros((3, 2, s1.shape[1]), order="F")
    out[:, 0] = s1
    out[:, 1] = s2
    return out


@pytest.mark.parametrize("Gc,rn", [(8, 1), (4, 0)])
def test_near_field_exact(prm, Gc, rn):
    dln, mat = prm
    net = random_loop_network(96, seed=0xD5)
    h = 2000.0 / Gc
    ff = G.FarField(Gc, (0.0, 0.0, 0.0), h, rn)
    S = 2 * Gc - 1
    ff.set_operator(np.zeros((S ** 3, 60, 90), np.float32))
    g, pairs, exact = G.calcSegSegForceHybrid(dln, mat, net, ff, farDist=1e30)
    n = int(net.numSeg[0])
    cell = H.bin_cells(net, Gc, (0, 0, 0), h)
    assert np.array_equal(ff.debug(2, n), cell.astype(np.int32))
    # like with like: each near pair once, oriented by the kernel's sorted
    # order.  The near-field kernel is built with FMA contraction and shared,
    # refined hardware reciprocals (the hybrid is a 1e-3 approximation), so
    # it differs from the IEEE oracle within the pair formula's conditioning:
    # up to ~2.6e-7 per node here, ~5e-9 RMS.
    once = O.segseg_near_once(dln, mat, net, cell, Gc, rn, H.sorted_rank(net, Gc, (0, 0, 0), h))
    mx, rms = per_node_rel(g, once)
    print(f"near G={Gc} rn={rn}: pairs {pairs} exact {exact}; max {mx:.2e} rms {rms:.2e}", flush=True)
    assert pairs == exact and mx <= 1e-6 and rms <= 1e-8
    # against the row-oriented near field: the pair formula's conditioning
    o = O.segseg_near(dln, mat, net, cell, Gc, rn)
    assert per_node_rel(g, o)[1] <= 1e-7
    # unordered pairs of live segments whose cells are within rn
    c3 = np.stack([cell % Gc, (cell // Gc) % Gc, cell // (Gc * Gc)], 1)[cell >= 0]
    cheb = np.abs(c3[:, None, :] - c3[None, :, :]).max(2)
    assert pairs == int((np.triu(cheb <= rn, 1)).sum())
    # with the mixed rule the far part of the stencil runs in FP32
    gm, pm, em = G.calcSegSegForceHybrid(dln, mat, net, ff)
    assert pm == pairs and em <= pm
    tot = o + _self(dln, mat, net)
    assert per_node_rel(gm + _self(dln, mat, net), tot)[0] <= 2e-5


def test_far_field_pipeline(prm):
    dln, mat = prm
    Gc, rn = 4, 1
    h = 500.0
    net = random_loop_network(24, seed=0xD6)
    ff = G.FarField(Gc, (0.0, 0.0, 0.0), h, rn)
    S = 2 * Gc - 1
    rng = np.random.default_rng(4)
    W = (rng.standard_normal((S ** 3, 60, 90)) * 1e-3).astype(np.float32)
    ff.set_operator(W)
    g, _, _ = G.calcSegSegForceHybrid(dln, mat, net, ff, farDist=1e30)
    M = ff.debug(0)
    Mr = H.moments(net, Gc, (0, 0, 0), h)
    assert np.abs(M - Mr).max() <= 1e-5 * np.abs(Mr).max()
    L = ff.debug(1)
    Lr = H.local_expansions(Mr, W.astype(np.float64), Gc, rn)
    assert np.abs(L - Lr).max() <= 1e-4 * np.abs(Lr).max()
    # far forces = hybrid - near
    cell = H.bin_cells(net, Gc, (0, 0, 0), h)
    near = O.segseg_near(dln, mat, net, cell, Gc, rn)
    far = H.far_forces(net, Lr, Gc, (0, 0, 0), h)
    diff = g - near
    assert np.abs(diff - far).max() <= 1e-4 * np.abs(far).max()


def _near_once(dln, mat, net, Gc, lo, h, rn):
    cell = H.bin_cells(net, Gc, lo, h)
    return O.segseg_near_once(dln, mat, net, cell, Gc, rn, H.sorted_rank(net, Gc, lo, h))


def test_hybrid_external_segments(prm):
    """External segments take no part (reference: seg-seg skips them)."""
    dln, mat = prm
    net = random_loop_network(48, seed=0xD7, ext_fraction=0.15)
    assert net.extSeg[:int(net.numSeg[0])].any()
    Gc, h, rn = 8, 250.0, 1
    ff = G.FarField(Gc, (0.0, 0.0, 0.0), h, rn)
    ff.set_operator(np.zeros(((2 * Gc - 1) ** 3, 60, 90), np.float32))
    g, pairs, exact = G.calcSegSegForceHybrid(dln, mat, net, ff, farDist=1e30)
    ref = _near_once(dln, mat, net, Gc, (0, 0, 0), h, rn)
    ext = net.extSeg[:int(net.numSeg[0])].astype(bool)
    assert not g[:, :, ext].any()
    assert per_node_rel(g, ref)[0] <= 1e-6


def test_hybrid_grid_smaller_than_network(prm):
    """Segments outside the grid are clamped into its border cells; the near
    field is still exactly the pairs of the clamped cells."""
    dln, mat = prm
    net = random_loop_network(64, seed=0xD8)
    Gc, lo, h, rn = 6, (500.0, 500.0, 500.0), 170.0, 1
    ff = G.FarField(Gc, lo, h, rn)
    ff.set_operator(np.zeros(((2 * Gc - 1) ** 3, 60, 90), np.float32))
    g, pairs, _ = G.calcSegSegForceHybrid(dln, mat, net, ff, farDist=1e30)
    assert np.array_equal(ff.debug(2, int(net.numSeg[0])), H.bin_cells(net, Gc, lo, h).astype(np.int32))
    assert per_node_rel(g, _near_once(dln, mat, net, Gc, lo, h, rn))[0] <= 1e-6


def test_hybrid_resident_matches_one_shot_and_repeats(prm):
    """The resident path (ddd_net_set_opts) gives the one-shot result bit for
    bit, and repeated evaluations are identical (deterministic sums)."""
    dln, mat = prm
    net = random_loop_network(96, seed=0xD9)
    Gc, h, rn = 8, 250.0, 1
    ff = G.FarField(Gc, (0.0, 0.0, 0.0), h, rn)
    rng = np.random.default_rng(9)
    ff.set_operator((rng.standard_normal(((2 * Gc - 1) ** 3, 60, 90)) * 1e-4).astype(np.float32))
    one, _, _ = G.calcSegSegForceHybrid(dln, mat, net, ff)
    res = G.ResidentNetwork(dln, mat, net)
    res.set_mode("hybrid", farfield=ff)
    res.seg_force()
    a = res.download()
    res.seg_force()
    b = res.download()
    s1, s2 = O.calcSelfForce(dln, mat, net)
    self_ = np.zeros_like(one)
    self_[:, 0], self_[:, 1] = s1, s2
    assert np.array_equal(a, b)
    assert np.allclose(a - self_, one, rtol=0, atol=1e-12 * np.abs(one).max())
    res.set_mode("fp64")


def test_near_parallel_list_growth(prm):
    """A pile-up of same-slip-system loops: more near-parallel pairs than the
    near-parallel list first holds (65,536), so the first evaluation enlarges
    the list and runs the tiles again.  Both evaluations agree bit for bit,
    and with the oracle's near pairs as the other tests do."""
    dln, mat = prm
    net = clustered_loop_network(numLoops=64, clusters=1, cluster_frac=1.0, sigma=60.0, seed=0xDA)
    n = int(net.numSeg[0])
    L = net.links[:, :n] - 1
    t = net.coord[:, L[1]] - net.coord[:, L[0]]
    t /= np.linalg.norm(t, axis=0)
    c = t.T @ t
    par = int(np.triu(1.0 - c * c <= 1.4901161193847656e-08, 1).sum())
    assert par > 1 << 16
    Gc, h, rn = 4, 500.0, 3  # every pair is near
    ff = G.FarField(Gc, (0.0, 0.0, 0.0), h, rn)
    ff.set_operator(np.zeros(((2 * Gc - 1) ** 3, 60, 90), np.float32))
    g1, pairs, exact = G.calcSegSegForceHybrid(dln, mat, net, ff, farDist=1e30)
    g2, _, _ = G.calcSegSegForceHybrid(dln, mat, net, ff, farDist=1e30)
    assert pairs == n * (n - 1) // 2 and exact == pairs
    assert np.array_equal(g1, g2)
    mx, rms = per_node_rel(g1, _near_once(dln, mat, net, Gc, (0.0, 0.0, 0.0), h, rn))
    print(f"pile-up: {n} segments, {par} near-parallel pairs; max {mx:.2e} rms {rms:.2e}", flush=True)
    assert mx <= 1e-6


@pytest.mark.parametrize("case", ["empty", "one", "parallel_pair", "one_loop"])
def test_hybrid_tiny_networks(prm, case):
    """Networks below one block: no segment, one segment, two parallel
    segments (the near-parallel list with a single entry) and one 32-sided
    loop (one tile, its 16 antiparallel pairs).  One-shot and resident
    evaluations agree with the oracle's near pairs; nothing is added below
    two segments."""
    dln, mat = prm
    b = np.array([1.0, 1.0, 1.0]) / np.sqrt(3.0)
    if case == "one_loop":
        net = random_loop_network(1, seed=0xDB)
    else:
        nseg = {"empty": 0, "one": 1, "parallel_pair": 2}[case]
        coord = np.array([[900.0, 1100.0, 900.0, 1100.0], [1000.0, 1000.0, 1040.0, 1040.0],
                          [1000.0, 1000.0, 1000.0, 1000.0]])[:, :2 * nseg]
        links = np.array([[1, 3], [2, 4]], dtype=np.int64)[:, :nseg]
        net = DislocationNetwork.from_arrays(links, np.tile(b[:, None], (1, nseg)), coord,
                                             [intMobDln] * (2 * nseg))
    n = int(net.numSeg[0])
    Gc, h, rn = 8, 250.0, 1
    ff = G.FarField(Gc, (0.0, 0.0, 0.0), h, rn)
    ff.set_operator(np.zeros(((2 * Gc - 1) ** 3, 60, 90), np.float32))
    g, pairs, exact = G.calcSegSegForceHybrid(dln, mat, net, ff, farDist=1e30)
    assert g.shape == (3, 2, n) and pairs == n * (n - 1) // 2 and exact == pairs
    res = G.ResidentNetwork(dln, mat, net)
    res.set_mode("hybrid", farfield=ff)
    res.seg_force()
    s1, s2 = O.calcSelfForce(dln, mat, net)
    self_ = np.zeros((3, 2, n), order="F")
    self_[:, 0], self_[:, 1] = s1, s2
    assert np.allclose(res.download() - self_, 0.0 if n < 2 else g, rtol=0,
                       atol=1e-12 * max(1.0, np.abs(g).max() if n else 1.0))
    res.set_mode("fp64")
    if n < 2:
        assert not g.any()
        return
    ref = _near_once(dln, mat, net, Gc, (0.0, 0.0, 0.0), h, rn)
    assert np.abs(ref).max() > 0
    assert per_node_rel(g, ref)[0] <= 1e-6
