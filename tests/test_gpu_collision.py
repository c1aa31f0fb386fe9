"""GPU parity of detectCollision (ddd_detect_collision:
DislocationTopologyCollision.jl:1-222) through the C-ABI: every check of the
reference's TopologyTest.jl:21-88 exactly, and the oracle's exact result on
random networks (the kernels round every distance, rate and 1/distance as
the oracle does, and the reductions keep the reference's first-in-loop-order
rule on ties), with and without skipped pairs."""
import numpy as np
import pytest

import pyoracle as O
import refnets as R
from dddhip import forces as G
from dddhip.generate import random_loop_network
from dddhip.types import DislocationNetwork, DislocationParameters, intFixDln

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("k", range(21))
def test_golden(k):
    c = R.GOLD["collision_cases"]["data"]["checks"][k]
    net = R.collision_network(c["net"])
    got = G.detectCollision(R.collision_params(), net, [tuple(p) for p in c["skip"]])
    exp = c["expect"]
    assert got[0] == exp[0] and got[1] == exp[1]
    assert list(got[2:8]) == [int(v) for v in exp[2:8]]
    assert got[8] == exp[8] and got[9] == exp[9]


def _moving(numLoops, seed, fixed_every=0):
    net = random_loop_network(numLoops, seed=seed)
    n = int(net.numNode[0])
    if fixed_every:
        lab = net.label.copy()
        lab[:n:fixed_every] = intFixDln
        net = DislocationNetwork._assemble(net.links, net.slipPlane, net.bVec, net.coord, lab, n,
                                           int(net.numSeg[0]), net.maxConnect)
    net.nodeVel[:, :n] = np.random.default_rng(seed).standard_normal((3, n))
    return net


@pytest.mark.parametrize("loops,dist,fixed", [(64, 157.235, 0), (64, 30.0, 3), (16, 60.0, 0)])
def test_random_networks_match_oracle(loops, dist, fixed):
    dln = DislocationParameters(coreRad=90.78, collisionDist=dist)
    net = _moving(loops, 0xC5 + loops, fixed)
    _log(f"collision {loops} loops, distance {dist}")
    assert G.detectCollision(dln, net) == O.detectCollision(dln, net)
    # no hinge: skip every hinge pair, so the unconnected pairs decide
    hinges = [(int(net.connectivity[1, i]), int(net.connectivity[3, i]))
              for i in range(int(net.numNode[0]))]
    g = G.detectCollision(dln, net, hinges)
    assert g == O.detectCollision(dln, net, hinges)
    assert g[1] in ("twoLine", "null")
    if g[0]:
        # and with the winner skipped too, the next one
        more = hinges + [(g[6], g[7])]
        assert G.detectCollision(dln, net, more) == O.detectCollision(dln, net, more)


def test_larger_network_matches_oracle():
    """8,192 segments, a collision distance below the hinge distances of
    32-gons, so that unconnected pairs decide."""
    dln = DislocationParameters(coreRad=90.78, collisionDist=3.0)
    net = _moving(256, 0xCA)
    _log("collision 8192 segments")
    g = G.detectCollision(dln, net)
    assert g == O.detectCollision(dln, net)


def test_no_collision_returns_last_evaluated_L():
    """Far apart: nothing collides; L1, L2 are those of the last pair the
    scan evaluated, as the reference returns them."""
    dln = DislocationParameters(coreRad=90.78, collisionDist=1e-3)
    net = _moving(32, 0xC9)
    g = G.detectCollision(dln, net)
    assert g[0] is False and g[1] == "null" and g == O.detectCollision(dln, net)


@pytest.mark.parametrize("case", range(2))
def test_junction_hinge_carry(case):
    """A node with three links: the reference's hinge loop compares each
    connection k with the longest of connections j..k-1 (its swap assigns
    the j loop's link variables); the hand-worked results are in
    tests/test_oracle_collision.py."""
    from test_oracle_collision import JUNCTION_CASES, junction_network, junction_params
    a_pos, a_vel, dist, expect = JUNCTION_CASES[case]
    _log(f"collision junction case {case}")
    net = junction_network(a_pos, a_vel)
    got = G.detectCollision(junction_params(dist), net)
    assert got[:8] == expect[:8]
    assert got[8] == expect[8] and got[9] == expect[9]
    assert got == O.detectCollision(junction_params(dist), net)


def test_last_pair_behind_fixed_links():
    """Links at the end of the array whose nodes are fixed are never
    evaluated; the last evaluated pair lies before them."""
    dln = DislocationParameters(coreRad=90.78, collisionDist=1e-3)
    net = _moving(64, 0xCB)
    n = int(net.numNode[0])
    lab = net.label.copy()
    lab[n - 640:n] = intFixDln  # the last 20 loops
    net = DislocationNetwork._assemble(net.links, net.slipPlane, net.bVec, net.coord, lab, n,
                                       int(net.numSeg[0]), net.maxConnect)
    net.nodeVel[:, :n] = np.random.default_rng(0xCB).standard_normal((3, n))
    _log("collision with fixed links at the end")
    g = G.detectCollision(dln, net)
    assert g[0] is False and g == O.detectCollision(dln, net)
