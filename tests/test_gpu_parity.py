"""GPU parity: the HIP path (through the C-ABI) against the reference's golden
vectors and against the CPU oracle on identical seeded networks.

Bar: reference vectors at the reference's own tolerance (Julia isapprox,
sqrt(eps) norm-wise); oracle comparisons at <= 1e-10 per-node relative error
(floor 1e-6 RMS, see metrics.per_node_rel).
"""
import numpy as np
import pytest

import pyoracle as O
import refnets as R
from dddhip import forces as G
from dddhip.generate import random_loop_network, face_grid_points, random_uhat
from dddhip.types import RegularCuboidMesh
from metrics import per_node_rel, isapprox, normwise

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def prm():
    return R.params()


# ---- reference golden vectors ------------------------------------------------

def test_golden_self_force(prm):
    dln, mat = prm
    net = R.pentagons(10.0)
    s1, s2 = G.calcSelfForce(dln, mat, net)
    assert isapprox(s1, R.g("self_f1").T) and isapprox(s2, R.g("self_f2").T)
    o1, o2 = O.calcSelfForce(dln, mat, net)
    assert per_node_rel(s1, o1)[0] <= TOL and per_node_rel(s2, o2)[0] <= TOL


@pytest.mark.parametrize("name,builder", [
    ("coincident", lambda: R.pentagons(0.0)),
    ("offset20", lambda: R.pentagons(20.0)),
    ("hexagons", R.hexagons),
])
def test_golden_segseg(prm, name, builder):
    dln, mat = prm
    net = builder()
    _log(f"segseg golden {name}")
    f = G.calcSegSegForce(dln, mat, net)
    assert isapprox(f[:, 0, :], R.g(f"segseg_{name}_f1").T)
    assert isapprox(f[:, 1, :], R.g(f"segseg_{name}_f2").T)
    o = O.calcSegSegForce(dln, mat, net)
    assert per_node_rel(f, o)[0] <= TOL


def test_golden_pk(prm):
    m = R.mesh()
    fd = R.sparse_uhat(m, "pk_uHat_index", "pk_uHat_value")
    net = R.pk_network()
    pk = G.calcPKForce(m, fd, net)
    assert isapprox(pk, R.g("pk_fPKTest").T)
    # in-place form: half to each node (SegmentForcesTest.jl:169-172)
    net.segForce[:] = 0
    G.calcPKForce_(m, fd, net)
    n = int(net.numSeg[0])
    assert isapprox(pk, net.segForce[:, 0, :n] * 2)
    assert np.array_equal(net.segForce[:, 0, :n], net.segForce[:, 1, :n])
    for idx in (5, [5, 10, 9, 23], range(15, 33)):
        p = G.calcPKForce(m, fd, net, idx)
        ii = np.atleast_1d(np.asarray(list(idx) if isinstance(idx, range) else idx)) - 1
        assert isapprox(p, R.g("pk_fPKTest")[ii].T)
    # outside the mesh the FE stress still interpolates from the clamped
    # element; the reference's check (:552-555) is that PK is zero when uHat
    # is zero there, which the zero-uHat mobility test covers.


def test_golden_sigma_hat():
    """calc_σHat through ddd_sigma_hat against SegmentForcesTest.jl:636-650
    (uHat :617-634), and against the oracle at points across the mesh,
    including points outside it (the clamped element) and on element faces."""
    m = R.mesh()
    fd = R.sparse_uhat(m, "sigmahat_uHat_index", "sigmahat_uHat_value")
    s1 = G.calc_σHat(m, fd, [1800.0, 1400.0, 200.0])
    s2 = G.calc_σHat(m, fd, [1893.0, 1356.0, 345.0])
    assert s1.shape == (3, 3)
    assert isapprox(s1, R.g("sigmahat_1"))
    assert isapprox(s2, R.g("sigmahat_2"))
    rng = np.random.default_rng(0x5E)
    faces = np.array([[0.0, 800.0, 2000.0], [400.0, 0.0, 2000.0], [2000.0, 1200.0, 0.0]])
    pts = np.asfortranarray(np.concatenate([rng.uniform(-200.0, 2200.0, (3, 61)), faces], axis=1))
    S = G.calc_σHat(m, fd, pts)
    assert S.shape == (3, 3, pts.shape[1])
    for p in range(pts.shape[1]):
        o = O.calc_σHat(m, fd, pts[:, p])
        assert np.array_equal(S[:, :, p], S[:, :, p].T)
        assert normwise(S[:, :, p], o) <= TOL, p


@pytest.mark.parametrize("b,key", [([1, 0, 0], "sigmatilde_edge"), ([0, 0, 1], "sigmatilde_screw")])
def test_golden_sigma_tilde(b, key):
    pts, dln, mat, net = R.sigma_segment(b)
    s = G.calc_σTilde(pts, dln, mat, net)
    t = R.g(key).copy()
    t[:, [4, 5]] = t[:, [5, 4]]  # the test matrices store (..., yz, xz)
    assert isapprox(s.T, t)
    σ = np.zeros((6, pts.shape[1]), order="F")
    G.calc_σTilde_(σ, pts, dln, mat, net)
    assert np.array_equal(σ, s)
    o = O.calc_σTilde(pts, dln, mat, net)
    assert normwise(s, o) <= TOL


def test_golden_node_force(prm):
    """Whole path + gather against MobilityTest.jl:78-90 (PK = 0)."""
    dln, mat = prm
    net = R.mobility_network()
    m = R.mesh()
    fd = R.ForceDisplacement(np.zeros(3 * m.numNode))
    G.calcSegForce_(dln, mat, m, fd, net)
    nf = G.calcNodeForce_(dln, mat, net)
    assert isapprox(nf, R.g("mobility_testForce").T)
    onet = R.mobility_network()
    sf = O.calcSegForce(dln, mat, m, fd, onet)
    onet.segForce[:, :, :sf.shape[2]] = sf
    assert per_node_rel(nf, O.nodeForce(mat, onet))[0] <= TOL


# ---- seeded random networks against the oracle -------------------------------

@pytest.fixture(scope="module")
def netA():
    # 20 loops x 32 segments, seed 0xA0
    return random_loop_network(20, seed=0xA0)


def test_config_a_segseg(prm, netA):
    dln, mat = prm
    _log("config A seg-seg")
    g = G.calcSegSegForce(dln, mat, netA)
    o = O.calcSegSegForce(dln, mat, netA)
    mx, rms = per_node_rel(g, o)
    _log(f"config A per-node rel err max {mx:.3e} rms {rms:.3e}")
    assert mx <= TOL


def test_config_a_total_with_pk(prm, netA):
    dln, mat = prm
    mesh = RegularCuboidMesh(dx=2000.0, dy=2000.0, dz=2000.0, mx=5, my=5, mz=5, μ=mat.μ, ν=mat.ν)
    fd = random_uhat(mesh, seed=0xA1)
    g = G.calcSegForce(dln, mat, mesh, fd, netA)
    o = O.calcSegForce(dln, mat, mesh, fd, netA)
    assert per_node_rel(g, o)[0] <= TOL
    # in-place on a buffer longer than numSeg (default memBuffer), then gather
    net = random_loop_network(20, seed=0xA0, memBuffer=None)
    assert net.segForce.shape[2] > int(net.numSeg[0])
    net.segForce[:] = 7.0  # calcSegForce! must zero the live range first
    G.calcSegForce_(dln, mat, mesh, fd, net)
    n = int(net.numSeg[0])
    assert per_node_rel(net.segForce[:, :, :n], o)[0] <= TOL
    assert np.all(net.segForce[:, :, n:] == 7.0)
    nf = G.calcNodeForce_(dln, mat, net)
    onet = random_loop_network(20, seed=0xA0, memBuffer=None)
    onet.segForce[:, :, :n] = o
    assert per_node_rel(nf, O.nodeForce(mat, onet))[0] <= TOL


def test_idx_rows(prm, netA):
    dln, mat = prm
    idx = [5, 17, 17, 600, 1, 640, 333]  # duplicates allowed, as rand(1:n, m) gives
    g = G.calcSegSegForce(dln, mat, netA, idx)
    o = O.calcSegSegForce(dln, mat, netA, idx)
    assert per_node_rel(g, o)[0] <= TOL
    # Rows put the row segment first in every pair, the full triangle puts
    # the lower index first; the two orders agree to the conditioning of the
    # pair formula, which the reference checks with isapprox
    # (SegmentForcesTest.jl:369-380).
    full = O.calcSegSegForce(dln, mat, netA)
    assert isapprox(g, full[:, :, np.asarray(idx) - 1])
    # in-place rows add into columns idx
    net = random_loop_network(20, seed=0xA0)
    G.calcSegSegForce_(dln, mat, net, 42)
    assert per_node_rel(net.segForce[:, :, 41], O.calcSegSegForce(dln, mat, net, 42))[0] <= TOL
    assert isapprox(net.segForce[:, :, 41], full[:, :, 41])
    assert np.count_nonzero(net.segForce[:, :, :41]) == 0
    s1, s2 = G.calcSelfForce(dln, mat, netA, range(3, 9))
    o1, o2 = O.calcSelfForce(dln, mat, netA, range(3, 9))
    assert per_node_rel(s1, o1)[0] <= TOL and per_node_rel(s2, o2)[0] <= TOL
    t = G.calcSegForce(dln, mat, None, None, netA, idx)
    ot = O.calcSegSegForce(dln, mat, netA, idx)
    a, b = O.calcSelfForce(dln, mat, netA, idx)
    ot[:, 0, :] += a
    ot[:, 1, :] += b
    assert per_node_rel(t, ot)[0] <= TOL


def test_idx_rows_chunked(prm, netA):
    """Rows split over column chunks (one row: 64 chunks at 16,384 segments;
    100 rows: 41 chunks) and, with 4,096 rows or more, one chunk per row
    added straight into the output; duplicates as rand(1:n, m) gives."""
    dln, mat = prm
    net = random_loop_network(512, seed=0xB0)
    n = int(net.numSeg[0])
    rng = np.random.default_rng(3)
    for m in (1, 100):
        idx = rng.integers(1, n + 1, m)
        _log(f"idx rows: {m} of {n} segments")
        g = G.calcSegSegForce(dln, mat, net, idx)
        o = O.calcSegSegForce(dln, mat, net, idx)
        assert per_node_rel(g, o)[0] <= TOL
    idx = rng.integers(1, int(netA.numSeg[0]) + 1, 5000)
    g = G.calcSegSegForce(dln, mat, netA, idx)
    o = O.calcSegSegForce(dln, mat, netA, idx)
    assert per_node_rel(g, o)[0] <= TOL
    assert np.array_equal(g, G.calcSegSegForce(dln, mat, netA, idx))


def test_external_segments(prm):
    dln, mat = prm
    net = random_loop_network(12, seed=0xA2, ext_fraction=0.15)
    assert net.extSeg[:int(net.numSeg[0])].any()
    g = G.calcSegSegForce(dln, mat, net)
    o = O.calcSegSegForce(dln, mat, net)
    assert per_node_rel(g, o)[0] <= TOL
    ext = net.extSeg[:int(net.numSeg[0])]
    assert np.all(g[:, :, ext] == 0)


def test_deterministic(prm):
    dln, mat = prm
    net = random_loop_network(64, seed=0xA3)
    a = G.calcSegSegForce(dln, mat, net)
    b = G.calcSegSegForce(dln, mat, net)
    assert np.array_equal(a, b)


def test_edge_cases(prm):
    dln, mat = prm
    net = random_loop_network(1, numSides=3, seed=0xA4)  # 3 segments, one tile
    assert per_node_rel(G.calcSegSegForce(dln, mat, net), O.calcSegSegForce(dln, mat, net))[0] <= TOL
    # 65 segments: one full tile plus a ragged one
    net = random_loop_network(13, numSides=5, seed=0xA5)
    assert per_node_rel(G.calcSegSegForce(dln, mat, net), O.calcSegSegForce(dln, mat, net))[0] <= TOL


def test_degenerate_inputs(prm):
    """Inputs at the edges of the domain: an empty network (no segments); idx
    entries outside 1:numSeg (an error, and the in-place form writes
    nothing); a zero-length segment and a NaN coordinate (the reference
    divides by the length and lets NaN through, so every segment that
    interacts with the bad one gets NaN: the NaN entries must be the
    oracle's); 64 and 129 segments (whole tiles, and two tiles plus a
    one-segment ragged one)."""
    from dddhip.types import DislocationNetwork
    dln, mat = prm
    empty = DislocationNetwork.from_arrays(np.zeros((2, 0), dtype=np.int64), np.zeros((3, 0)),
                                           np.zeros((3, 0)), [])
    _log("empty network")
    assert G.calcSegSegForce(dln, mat, empty).shape == (3, 2, 0)
    assert G.calcSegForce(dln, mat, None, None, empty).shape == (3, 2, 0)
    pts = face_grid_points(4)
    s = G.calc_σTilde(pts, dln, mat, empty)
    assert s.shape == (6, pts.shape[1]) and np.all(s == 0)

    net = random_loop_network(4, seed=0xA7)
    n = int(net.numSeg[0])
    _log("idx outside 1:numSeg")
    for bad in ([0], [3, n + 1]):
        with pytest.raises(RuntimeError, match="idx entry outside"):
            G.calcSegSegForce(dln, mat, net, bad)
    net.segForce[...] = 3.0
    with pytest.raises(RuntimeError, match="idx entry outside"):
        G.calcSegSegForce_(dln, mat, net, [n + 1])
    assert np.all(net.segForce == 3.0)

    for what in ("zero-length", "nan"):
        net = random_loop_network(4, seed=0xA8)
        si = net.segIdx[5]
        if what == "zero-length":
            net.coord[:, si[2] - 1] = net.coord[:, si[1] - 1]
        else:
            net.coord[1, si[1] - 1] = np.nan
        _log(f"{what} segment")
        g = G.calcSegSegForce(dln, mat, net)
        o = O.calcSegSegForce(dln, mat, net)
        # a zero-length segment makes its own forces NaN (the other
        # segments see a degenerate projection with finite forces); a NaN
        # coordinate reaches every pair
        assert np.array_equal(np.isnan(g), np.isnan(o)) and np.isnan(g).any()
        fin = np.isfinite(o)
        if fin.any():  # the sums run in another order: equal to rounding
            assert np.max(np.abs(g[fin] - o[fin])) <= TOL * np.max(np.abs(o[fin]))

    for loops, sides in ((2, 32), (3, 43)):
        net = random_loop_network(loops, numSides=sides, seed=0xA9)
        assert int(net.numSeg[0]) == loops * sides
        _log(f"{loops * sides} segments")
        g = G.calcSegSegForce(dln, mat, net)
        assert per_node_rel(g, O.calcSegSegForce(dln, mat, net))[0] <= TOL


def test_tiles_outside_the_ranged_range(prm):
    """The tile kernel's ranged division and square root apply only where
    every coordinate and Burgers component is zero or of magnitude in
    [2^-300, 2^60] and a^2 lies in [2^-120, 2^120]; other tiles go to the
    general IEEE sequences.  Parity with the oracle when some tiles (one loop
    moved to x + 2^61, a node coordinate of 1e-200) or all tiles (the whole
    network and the core radius scaled by 2^62) fall outside."""
    from dddhip.types import DislocationParameters
    dln, mat = prm
    for what in ("one loop at 2^61", "a coordinate of 1e-200", "all scaled by 2^62"):
        net = random_loop_network(8, seed=0xAB)
        d = dln
        if what == "one loop at 2^61":
            nodes = np.unique(net.segIdx[:32, 1:].ravel()) - 1
            net.coord[0, nodes] += 2.0 ** 61
        elif what == "a coordinate of 1e-200":
            net.coord[2, net.segIdx[40, 1] - 1] = 1e-200
        else:
            net.coord[:, :] *= 2.0 ** 62
            d = DislocationParameters(coreRad=dln.coreRad * 2.0 ** 62, coreEnergy=dln.coreEnergy)
        _log(what)
        g = G.calcSegSegForce(d, mat, net)
        o = O.calcSegSegForce(d, mat, net)
        assert np.all(np.isfinite(o))
        assert per_node_rel(g, o)[0] <= TOL


def test_sigma_tilde_random(prm):
    dln, mat = prm
    net = random_loop_network(20, seed=0xA6)
    pts = face_grid_points(16)
    g = G.calc_σTilde(pts, dln, mat, net)
    o = O.calc_σTilde(pts, dln, mat, net)
    mx, _ = per_node_rel(g[:3], o[:3])
    assert normwise(g, o) <= TOL
    assert np.max(np.abs(g - o)) <= TOL * np.max(np.abs(o))
    # idx subset and single point
    sub = G.calc_σTilde(pts, dln, mat, net, [3, 1, 200])
    assert normwise(sub, O.calc_σTilde(pts, dln, mat, net, [3, 1, 200])) <= TOL
    one = G.calc_σTilde(pts[:, 7], dln, mat, net)
    assert one.shape == (6,) and normwise(one, o[:, 7]) <= TOL


def test_medium_network(prm):
    """4,096 segments against the oracle's threaded restatement."""
    dln, mat = prm
    net = random_loop_network(128, seed=0xB1)
    _log("medium 4096 seg-seg")
    g = G.calcSegSegForce(dln, mat, net)
    o = O.calcSegSegForcePar(dln, mat, net, 16)
    mx, rms = per_node_rel(g, o)
    _log(f"medium per-node rel err max {mx:.3e} rms {rms:.3e}")
    assert mx <= TOL


def test_config_b(prm):
    """configs[1] size (16,384 segments, 1.34e8 pairs): the full GPU triangle
    against the oracle's threaded triangle, GPU rows against oracle rows, and
    the vanishing net force of closed loops."""
    dln, mat = prm
    net = random_loop_network(512, seed=0xB0)
    _log("config B seg-seg")
    g = G.calcSegSegForce(dln, mat, net)
    o = O.calcSegSegForcePar(dln, mat, net, 16)
    mx, rms = per_node_rel(g, o)
    _log(f"config B per-node rel err max {mx:.3e} rms {rms:.3e}")
    assert mx <= TOL
    rows = np.linspace(1, 16384, 48).astype(np.int64)
    orow = O.calcSegSegForce(dln, mat, net, rows)
    gr = G.calcSegSegForce(dln, mat, net, rows)
    assert per_node_rel(gr, orow)[0] <= TOL
    # Net force of pair interactions of closed loops vanishes.
    tot = g.reshape(3, -1, order="F").sum(axis=1)
    assert np.all(np.abs(tot) <= 1e-12 * np.abs(g).sum())


def test_pileup_near_parallel_pairs(prm):
    """A network of pile-ups (loops of one slip system, so their sides are
    parallel): near-parallel pairs are a large share (listed per tile and
    evaluated by their own kernel), against the oracle's triangle at 1e-10
    per node, and bitwise repeatable."""
    from dddhip.generate import clustered_loop_network
    dln, mat = prm
    net = clustered_loop_network(numLoops=96, seed=0xE7, clusters=2, cluster_frac=0.9,
                                 sigma=120.0)
    n = int(net.numSeg[0])
    _log(f"pile-up network: {n} segments")
    g = G.calcSegSegForce(dln, mat, net)
    o = O.calcSegSegForcePar(dln, mat, net, 16)
    mx, rms = per_node_rel(g, o)
    _log(f"pile-up per-node rel err max {mx:.3e} rms {rms:.3e}")
    assert mx <= TOL
    assert np.array_equal(g, G.calcSegSegForce(dln, mat, net))
    # share of near-parallel pairs (1 - c^2 <= sqrt(eps)) in this network
    si = net.segIdx[:n]
    d = net.coord[:, si[:, 2] - 1] - net.coord[:, si[:, 1] - 1]
    t = d / np.linalg.norm(d, axis=0)
    c = t.T @ t
    par = np.triu(1 - c * c <= np.sqrt(np.finfo(float).eps), 1).sum()
    _log(f"near-parallel pairs: {par} of {n * (n - 1) // 2}")
    assert par > 0.02 * n * (n - 1) // 2


def test_sigma_tilde_config_b(prm):
    """calc_σTilde of the 16,384-segment network at the 6 x 128^2 face points:
    every 48th point against the oracle (1e-10 per stress component of the
    largest), all points finite, and the full call consistent with a call on
    a point subset (points are independent)."""
    dln, mat = prm
    net = random_loop_network(512, seed=0xB0)
    pts = face_grid_points(128)
    _log("config B sigma tilde")
    g = G.calc_σTilde(pts, dln, mat, net)
    assert np.isfinite(g).all()
    sel = np.arange(0, pts.shape[1], 48)
    o = O.calc_σTilde(np.asfortranarray(pts[:, sel]), dln, mat, net)
    err = np.abs(g[:, sel] - o).max() / np.abs(o).max()
    _log(f"config B sigma tilde: {len(sel)} points, max rel err {err:.2e}")
    assert err <= TOL
    # a point subset is split into a different number of segment chunks, so
    # its sums associate differently: equal to rounding, not bit for bit
    part = G.calc_σTilde(np.asfortranarray(pts[:, 1000:3000]), dln, mat, net)
    assert np.abs(part - g[:, 1000:3000]).max() <= 1e-12 * np.abs(g).max()
