"""CPU tests of the oracle: it is pinned against the reference's golden vectors,
against independent mathematics (quadrature of the stress field, closed-loop
force balance, symmetry and scaling invariants) and its log/atan against
mpmath.  A plausible slip in the restatement (a dropped term, a sign, a
transposed operand) breaks at least one of these."""
import math

import mpmath as mp
import numpy as np
import pytest

import pyoracle as O
import refnets as R
from dddhip.generate import random_loop_network
from dddhip.types import DislocationNetwork, MaterialParameters
from metrics import isapprox, per_node_rel, normwise


@pytest.fixture(scope="module")
def prm():
    return R.params()


# ---- the reference's golden vectors ----------------------------------------

def test_golden_self(prm):
    dln, mat = prm
    s1, s2 = O.calcSelfForce(dln, mat, R.pentagons(10.0))
    assert isapprox(s1, R.g("self_f1").T) and isapprox(s2, R.g("self_f2").T)
    assert np.array_equal(s1, -s2)


@pytest.mark.parametrize("name,builder", [("coincident", lambda: R.pentagons(0.0)),
                                          ("offset20", lambda: R.pentagons(20.0)),
                                          ("hexagons", R.hexagons)])
def test_golden_segseg(prm, name, builder):
    dln, mat = prm
    net = builder()
    f = O.calcSegSegForce(dln, mat, net)
    assert isapprox(f[:, 0, :], R.g(f"segseg_{name}_f1").T)
    assert isapprox(f[:, 1, :], R.g(f"segseg_{name}_f2").T)
    # the threaded design gives the same answer (SegmentForcesTest.jl:323-324)
    assert isapprox(O.calcSegSegForcePar(dln, mat, net, 3), f)


def test_golden_pk_and_sigma_hat():
    m = R.mesh()
    fd = R.sparse_uhat(m, "pk_uHat_index", "pk_uHat_value")
    net = R.pk_network()
    assert isapprox(O.calcPKForce(m, fd, net), R.g("pk_fPKTest").T)
    assert isapprox(O.calcPKForce(m, fd, net, [5, 10, 9, 23]), R.g("pk_fPKTest")[[4, 9, 8, 22]].T)
    fd2 = R.sparse_uhat(m, "sigmahat_uHat_index", "sigmahat_uHat_value")
    assert isapprox(O.calc_σHat(m, fd2, [1800.0, 1400.0, 200.0]), R.g("sigmahat_1"))
    assert isapprox(O.calc_σHat(m, fd2, [1893.0, 1356.0, 345.0]), R.g("sigmahat_2"))


@pytest.mark.parametrize("b,key", [([1, 0, 0], "sigmatilde_edge"), ([0, 0, 1], "sigmatilde_screw")])
def test_golden_sigma_tilde(b, key):
    pts, dln, mat, net = R.sigma_segment(b)
    t = R.g(key).copy()
    t[:, [4, 5]] = t[:, [5, 4]]  # the reference's test matrices store (..., yz, xz)
    assert isapprox(O.calc_σTilde(pts, dln, mat, net).T, t)


def test_golden_node_force(prm):
    dln, mat = prm
    net = R.mobility_network()
    m = R.mesh()
    sf = O.calcSegForce(dln, mat, m, R.ForceDisplacement(np.zeros(3 * m.numNode)), net)
    net.segForce[:, :, :sf.shape[2]] = sf
    assert isapprox(O.nodeForce(mat, net), R.g("mobility_testForce").T)


# ---- log / atan --------------------------------------------------------------

def _ulps(a, ref):
    return abs(a - ref) / np.spacing(abs(ref)) if ref != 0 else abs(a)


def test_libm_against_mpmath():
    mp.mp.prec = 120
    rng = np.random.default_rng(0)
    xs = np.concatenate([np.exp(rng.uniform(-700, 700, 3000)), rng.uniform(0.5, 2.0, 3000),
                         1 + rng.normal(size=500) * 1e-7, [1.0, 2.0, 0.5, 1e-310, 5e-324]])
    worst = max(_ulps(O.libm_log(x), float(mp.log(mp.mpf(x)))) for x in xs)
    assert worst <= 1.0
    ys = np.concatenate([rng.normal(size=3000) * 10.0 ** rng.uniform(-12, 12, 3000),
                         rng.uniform(-3, 3, 3000), [0.0, 0.4375, 0.6875, 1.1875, 2.4375, 1e30]])
    worst = max(_ulps(O.libm_atan(y), float(mp.atan(mp.mpf(y)))) for y in ys)
    assert worst <= 1.0
    assert O.libm_log(0.0) == -math.inf and math.isnan(O.libm_log(-1.0))
    assert O.libm_log(math.inf) == math.inf and math.isnan(O.libm_atan(math.nan))
    assert O.libm_atan(-2.0) == -O.libm_atan(2.0)


def test_libm_choice_sensitivity(prm):
    """Changing log/atan implementation only moves results at the level of
    the formula's own conditioning (documented, not a parity bar)."""
    dln, mat = prm
    net = random_loop_network(8, seed=0xA0)
    a = O.calcSegSegForce(dln, mat, net)
    O.set_libm(True)
    try:
        b = O.calcSegSegForce(dln, mat, net)
    finally:
        O.set_libm(False)
    assert per_node_rel(a, b)[0] <= 1e-7


# ---- independent mathematics ------------------------------------------------

def _sigma_of_segment(dln, mat, b, n1, n2, pts):
    net = DislocationNetwork.from_arrays(np.array([[1], [2]]), np.asarray(b, float).reshape(3, 1),
                                         np.stack([n1, n2], 1), [0, 0])
    return O.calc_σTilde(pts, dln, mat, net, nthreads=1)


def _quad_nodal_force(dln, mat, b1, a1, a2, b2, c1, c2, n=96):
    """Nodal forces on segment (a1, a2) from segment (c1, c2) by Gauss-Legendre
    quadrature of the Peach-Koehler force density of seg 2's non-singular
    stress against the linear shape functions of seg 1."""
    x, w = np.polynomial.legendre.leggauss(n)
    u, w = (x + 1) / 2, w / 2
    pts = a1[:, None] + (a2 - a1)[:, None] * u[None, :]
    s = _sigma_of_segment(dln, mat, b2, c1, c2, pts)
    L = np.linalg.norm(a2 - a1)
    t = (a2 - a1) / L
    F1, F2 = np.zeros(3), np.zeros(3)
    for k in range(n):
        S = np.array([[s[0, k], s[3, k], s[4, k]], [s[3, k], s[1, k], s[5, k]],
                      [s[4, k], s[5, k], s[2, k]]])
        f = np.cross(S @ b1, t) * L
        F1 += w[k] * (1 - u[k]) * f
        F2 += w[k] * u[k] * f
    return F1, F2


def test_pair_against_quadrature(prm):
    dln, mat = prm
    K = O.seg_consts(dln, mat)
    rng = np.random.default_rng(5)
    for _ in range(8):
        b1, b2 = rng.normal(size=3), rng.normal(size=3)
        a1 = rng.normal(size=3) * 100
        a2 = a1 + rng.normal(size=3) * 50
        c1 = rng.normal(size=3) * 100 + 200
        c2 = c1 + rng.normal(size=3) * 50
        F, par = O.pair(b1, a1, a2, b2, c1, c2, K)
        assert not par
        q1, q2 = _quad_nodal_force(dln, mat, b1, a1, a2, b2, c1, c2)
        assert np.abs(F[0] - q1).max() <= 1e-8 * np.abs(q1).max()
        assert np.abs(F[1] - q2).max() <= 1e-8 * np.abs(q2).max()
        r1, r2 = _quad_nodal_force(dln, mat, b2, c1, c2, b1, a1, a2)
        assert np.abs(F[2] - r1).max() <= 1e-8 * np.abs(r1).max()
        assert np.abs(F[3] - r2).max() <= 1e-8 * np.abs(r2).max()


def test_parallel_branch_against_quadrature(prm):
    """The near-parallel branch tilts one segment by the critical angle
    (~1.2e-4 rad): it agrees with the exact integral to that order."""
    dln, mat = prm
    K = O.seg_consts(dln, mat)
    rng = np.random.default_rng(6)
    for sgn in (1.0, -1.0):
        b1, b2 = rng.normal(size=3), rng.normal(size=3)
        t = rng.normal(size=3)
        t /= np.linalg.norm(t)
        a1 = rng.normal(size=3) * 50
        a2 = a1 + 40 * t
        off = np.cross(t, rng.normal(size=3)) * 2
        c1 = a1 + off + 10 * t
        c2 = c1 + sgn * 60 * t
        F, par = O.pair(b1, a1, a2, b2, c1, c2, K)
        assert par
        q1, q2 = _quad_nodal_force(dln, mat, b1, a1, a2, b2, c1, c2, n=200)
        scale = max(np.abs(q1).max(), np.abs(q2).max())
        assert np.abs(F[0] - q1).max() <= 1e-3 * scale
        assert np.abs(F[1] - q2).max() <= 1e-3 * scale


def test_pair_invariants(prm):
    dln, mat = prm
    K = O.seg_consts(dln, mat)
    rng = np.random.default_rng(7)
    for _ in range(20):
        b1, b2 = rng.normal(size=3), rng.normal(size=3)
        a1 = rng.normal(size=3) * 100
        a2 = a1 + rng.normal(size=3) * 40
        c1 = rng.normal(size=3) * 100 + 150
        c2 = c1 + rng.normal(size=3) * 40
        F, _ = O.pair(b1, a1, a2, b2, c1, c2, K)
        # The generic formula loses ~(distance / length)^2 ulps to cancellation,
        # so invariants hold to 1e-7 of the force scale; a wrong term or sign
        # breaks them at O(1).
        sc = 1e2 * np.abs(F).max()
        # swapping the segments swaps the outputs
        Fs, _ = O.pair(b2, c1, c2, b1, a1, a2, K)
        assert np.abs(Fs[[2, 3, 0, 1]] - F).max() <= 1e-9 * sc
        # rigid translation
        d = rng.normal(size=3) * 300
        Ft, _ = O.pair(b1, a1 + d, a2 + d, b2, c1 + d, c2 + d, K)
        assert np.abs(Ft - F).max() <= 1e-9 * sc
        # rotation covariance F(Rx, Rb) = R F
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        Fr, _ = O.pair(Q @ b1, Q @ a1, Q @ a2, Q @ b2, Q @ c1, Q @ c2, K)
        assert np.abs(Fr - F @ Q.T).max() <= 1e-9 * sc
        # bilinear in (b1, b2)
        F2, _ = O.pair(2.5 * b1, a1, a2, -0.5 * b2, c1, c2, K)
        assert np.abs(F2 + 1.25 * F).max() <= 1e-12 * sc
        # segment-sum force balance of the pair, F1+F2 = -(F3+F4) only for
        # closed loops; here check scaling: F(s x, s a) = F(x, a)
        s = 3.0
        Ks = O.seg_consts(type(dln)(coreRad=dln.coreRad * s, coreEnergy=1.0), mat)
        Fsc, _ = O.pair(b1, s * a1, s * a2, b2, s * c1, s * c2, Ks)
        assert np.abs(Fsc - F).max() <= 1e-9 * sc
    # linear in mu at fixed nu
    mat2 = MaterialParameters(μ=3.0, ν=mat.ν)
    F, _ = O.pair(b1, a1, a2, b2, c1, c2, K)
    F3, _ = O.pair(b1, a1, a2, b2, c1, c2, O.seg_consts(dln, mat2))
    assert np.abs(F3 - 3 * F).max() <= 1e-12 * np.abs(F).max()


def test_closed_loop_force_balance(prm):
    """For closed loops the pair forces over all node ends sum to zero."""
    dln, mat = prm
    for net in (R.pentagons(20.0), R.hexagons(), random_loop_network(10, seed=0xC1)):
        f = O.calcSegSegForce(dln, mat, net)
        tot = f.reshape(3, -1, order="F").sum(axis=1)
        assert np.all(np.abs(tot) <= 1e-12 * np.abs(f).sum())


def test_sigma_tilde_linear_in_b_and_additive(prm):
    dln, mat = prm
    net = random_loop_network(3, seed=0xC2)
    pts = np.asfortranarray(np.random.default_rng(1).uniform(-200, 2200, (3, 40)))
    full = O.calc_σTilde(pts, dln, mat, net)
    n = int(net.numSeg[0])
    part = (O.calc_σTilde(pts, dln, mat, net, range(1, n // 2 + 1))
            + O.calc_σTilde(pts, dln, mat, net, range(n // 2 + 1, n + 1)))
    assert normwise(full, part) <= 1e-13
    net.bVec *= -2.0
    assert normwise(O.calc_σTilde(pts, dln, mat, net), -2.0 * full) <= 1e-14


def test_self_force_closed_form(prm):
    """Arsenlis et al. 2007 App. A (quoted at DislocationForces.jl:431-439)
    evaluated independently for a pure screw and a pure edge segment."""
    dln, mat = prm
    a, Ec, nu, mu = dln.coreRad, dln.coreEnergy, mat.ν, mat.μ
    L = 37.0
    La = math.sqrt(L * L + a * a)
    for kind in ("screw", "edge"):
        b = np.array([1.0, 0.0, 0.0]) if kind == "screw" else np.array([0.0, 1.0, 0.0])
        net = DislocationNetwork.from_arrays(np.array([[1], [2]]), b.reshape(3, 1),
                                             np.array([[0.0, L], [0.0, 0.0], [0.0, 0.0]]), [1, 1])
        _, f2 = O.calcSelfForce(dln, mat, net)
        bs = b[0]
        be2 = 1.0 - bs * bs
        tor = mu / (4 * math.pi) * bs * (nu / (1 - nu) * (math.log((La + L) / a)
                                          - 2 * (La - a) / L) - (La - a) ** 2 / (2 * La * L))
        tor += 2 * Ec * nu / (1 - nu) * bs
        lon = (bs * bs + be2 / (1 - nu)) * Ec
        expect = tor * (b - bs * np.array([1.0, 0, 0])) - lon * np.array([1.0, 0, 0])
        assert np.abs(f2[:, 0] - expect).max() <= 1e-14 * max(1.0, np.abs(expect).max())


def test_op_counts(prm):
    """Per-unit work figures used for the roofline."""
    dln, mat = prm
    K = O.seg_consts(dln, mat)
    rng = np.random.default_rng(2)
    v = [rng.normal(size=3) * s for s in (1, 100, 100, 1, 100, 100)]
    c, par = O.count_pair(*v, K)
    assert not par
    assert (c["flops"], c["div"], c["sqrt"], c["log"], c["atan"]) == (1466, 23, 10, 8, 4)
    s = O.count_sigma_tilde(O.sigma_consts(dln, mat), [1, 0, 0], [0, 0, -3], [0, 0, 0], [1, 2, 3])
    assert (s["flops"], s["div"], s["sqrt"]) == (291, 6, 3)
