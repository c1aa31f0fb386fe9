"""CPU tests of the oracle's collision detection: minimumDistance
(Misc.jl:314-367) against every case of MiscTest.jl:82-165 and
detectCollision (DislocationTopologyCollision.jl:1-222) against every check
of TopologyTest.jl:21-88, exactly (both return exact small numbers there),
plus brute force: the minimum distance of two segments sampled densely."""
import numpy as np
import pytest

import pyoracle as O
import refnets as R


@pytest.mark.parametrize("k", range(12))
def test_minimum_distance_golden(k):
    c = R.GOLD["min_distance_cases"]["data"][k]
    got = O.minimumDistance(*[c["in"][n] for n in ("x0", "x1", "y0", "y1", "vx0", "vx1", "vy0",
                                                   "vy1")])
    assert list(got) == c["expect"]


def test_minimum_distance_against_sampling():
    """Generic pairs: the reference solves for the closest points of the two
    infinite lines and clamps each parameter to [0, 1] on its own, so the
    result is the segments' true minimum distance when neither parameter
    was clamped, and otherwise the distance between the two clamped points
    (never below the true minimum); dDistSqDt is the time derivative of
    |r|^2 at those points (finite difference)."""
    rng = np.random.default_rng(2)
    s = np.linspace(0, 1, 401)
    inner = 0
    for _ in range(60):
        x0, x1, y0, y1, vx0, vx1, vy0, vy1 = rng.standard_normal((8, 3))
        d2, dd, L1, L2 = O.minimumDistance(x0, x1, y0, y1, vx0, vx1, vy0, vy1)
        P = x0[None] + s[:, None] * (x1 - x0)[None]
        Q = y0[None] + s[:, None] * (y1 - y0)[None]
        brute = ((P[:, None] - Q[None]) ** 2).sum(-1).min()
        pts = ((x0 + L1 * (x1 - x0)) - (y0 + L2 * (y1 - y0)))
        assert abs(d2 - pts @ pts) <= 1e-12 * (1 + d2)
        assert d2 >= brute - 1e-4 * (1 + brute)
        if 0 < L1 < 1 and 0 < L2 < 1:
            inner += 1
            assert d2 <= brute + 1e-12
        # stationarity of the unclamped solution: r is normal to both lines
        if 0 < L1 < 1 and 0 < L2 < 1:
            assert abs(pts @ (x1 - x0)) <= 1e-9 and abs(pts @ (y1 - y0)) <= 1e-9
        h = 1e-6
        r = lambda t: ((x0 + t * vx0) + L1 * ((x1 + t * vx1) - (x0 + t * vx0))
                       - (y0 + t * vy0) - L2 * ((y1 + t * vy1) - (y0 + t * vy0)))
        fd = ((r(h) ** 2).sum() - (r(-h) ** 2).sum()) / (2 * h)
        assert abs(fd - dd) <= 1e-6 * (1 + abs(dd))
    assert inner >= 5


@pytest.mark.parametrize("k", range(21))
def test_detect_collision_golden(k):
    c = R.GOLD["collision_cases"]["data"]["checks"][k]
    net = R.collision_network(c["net"])
    got = O.detectCollision(R.collision_params(), net, [tuple(p) for p in c["skip"]])
    exp = c["expect"]
    assert got[0] == exp[0] and got[1] == exp[1]
    assert list(got[2:8]) == [int(v) for v in exp[2:8]]
    assert got[8] == exp[8] and got[9] == exp[9]


def junction_network(a_pos, a_vel):
    """Node 1 at the origin joined to A (node 2, the shortest link, link 1),
    B = (0, 3, 0) (node 3, the longest, link 2) and C = (0, 0, 2) (node 4,
    link 3), in that connection order; every node mobile, only A moves."""
    from dddhip.types import DislocationNetwork, intMobDln
    links = np.array([[1, 1, 1], [2, 3, 4]])
    coord = np.array([[0.0, 0.0, 0.0], a_pos, [0.0, 3.0, 0.0], [0.0, 0.0, 2.0]]).T
    bVec = np.tile(np.array([[1.0], [1.0], [1.0]]) / np.sqrt(3.0), (1, 3))
    net = DislocationNetwork.from_arrays(links, bVec, coord, [intMobDln] * 4)
    net.nodeVel[:, 1] = a_vel
    return net


def junction_params(dist):
    from dddhip.types import DislocationParameters
    return DislocationParameters(coreRad=90.78, collisionDist=dist)


# The reference's hinge loop at node 1 (DislocationTopologyCollision.jl:24-126)
# worked by hand.  link1/connectedNode1/tVecN1 are the j loop's variables
# and the swap at :52-68 assigns them, so each k is compared with the longest
# of connections j..k-1:
#   j = 1 (A): k = 2 (B): |A|^2 < |B|^2, swap -> pair (B, A);
#              k = 3 (C): link 1 is now B, |B|^2 > |C|^2 -> pair (B, C);
#   j = 2 (B): k = 3 (C) -> pair (B, C), the last pair evaluated.
# Starting each k from connection j afresh would give (B, A), (C, A), (B, C).
# Every link shares node 1, so the two-line scan evaluates nothing, and L1, L2
# are those of (B, C): point C against segment 1 -> B, clamped at node 1.
JUNCTION_CASES = [
    # A next to segment 1 -> C only: (C, A) would collide, (B, A) is 0.51 away
    ((0.1, 0.0, 0.5), (-1.0, 0.0, 0.0), 0.2,
     (False, "null", 0, 0, 0, 0, 0, 0, 0.0, 0.0)),
    # A within 0.6 of both lines: (C, A) at 0.316 would win, (B, A) at 0.51 does
    ((0.1, 0.3, 0.5), (-1.0, -1.0, -1.0), 0.6,
     (True, "hinge", 1, 3, 2, 2, 2, 1, 0.0, 0.0)),
]


@pytest.mark.parametrize("case", range(len(JUNCTION_CASES)))
def test_detect_collision_junction_hinge_carry(case):
    a_pos, a_vel, dist, expect = JUNCTION_CASES[case]
    got = O.detectCollision(junction_params(dist), junction_network(a_pos, a_vel))
    assert got[:8] == expect[:8]
    assert got[8] == expect[8] and got[9] == expect[9]
