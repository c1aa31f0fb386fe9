"""CPU tests of the oracle's deriv!/integrate! (IntegrationProcessing.jl:13-168).

The reference's tests hold no trajectory, so these pin the restatement by
what the algorithm fixes in closed form: the step-size controller on a
network that cannot move (dt grows by maxchange per accepted trial until
maxiter trials were accepted, or stops at dtmax), the forward-Euler trial at
dtmin, the halving of rejected steps; and by physics: a glide loop under its
own line tension shrinks and stays in its glide plane (climb drag 1e10),
fixed nodes stay put.  The velocities themselves are the mobility's, pinned
in test_oracle_mobility.py."""
import math

import numpy as np
import pytest

import pyoracle as O
import refnets as R
from dddhip.generate import random_loop_network
from dddhip.types import (DislocationLoop, DislocationNetwork, IntegrationParameters,
                          IntegrationTime, SlipSystem, intFixDln, intMobDln)


@pytest.fixture(scope="module")
def prm():
    return R.params()


def _glide_loop(R0=400.0, sides=16, slip=1):
    lp = DislocationLoop(loopType="shear", numSides=sides, nodeSide=1, numLoops=1,
                         segLen=2 * R0 * math.sin(math.pi / sides) * np.ones(sides),
                         slipSystemIdx=slip, slipSystem=SlipSystem.bcc(),
                         label=[intMobDln] * sides, buffer=0.0,
                         range=np.array([[0.0, 2000.0]] * 3), dist=np.full((3, 1), 0.5))
    return DislocationNetwork.from_loops(lp, memBuffer=1)


def _fixed(net):
    n = int(net.numNode[0])
    lab = np.full(len(net.label), intFixDln)
    return DislocationNetwork._assemble(net.links, net.slipPlane, net.bVec, net.coord, lab, n,
                                        int(net.numSeg[0]), net.maxConnect)


def test_controller_grows_dt_when_nothing_moves(prm):
    """All nodes fixed: every trial has zero distance and error, so each is
    accepted with factor = maxchange; after maxiter acceptances the step
    returns dt0 maxchange^maxiter and time + that dt, coordinates unchanged."""
    dln, mat = prm
    net = _fixed(_glide_loop())
    c0 = net.coord.copy()
    ip = IntegrationParameters(dtmin=1e-6, dtmax=1e15, abstol=10.0, reltol=1.0, maxchange=1.2,
                               exponent=20.0, maxiter=7)
    dt, time, step = O.integrate_(ip, IntegrationTime(dt=100.0, time=5.0, step=3), dln, mat,
                                  None, None, net)
    want = 100.0
    for _ in range(7):
        want = min(want * 1.2, ip.dtmax)
    assert dt == want and time == 5.0 + want and step == 4
    assert np.array_equal(net.coord, c0)
    assert np.all(net.nodeVel[:, :int(net.numNode[0])] == 0)
    # dtmax reached on the first acceptance ends the step at dtmax
    ip2 = IntegrationParameters(dtmin=1e-6, dtmax=110.0, abstol=10.0, reltol=1.0, maxiter=7)
    dt, time, step = O.integrate_(ip2, IntegrationTime(dt=100.0), dln, mat, None, None, net)
    assert (dt, time, step) == (110.0, 110.0, 1)


def test_trial_at_dtmin_is_forward_euler(prm):
    """dt <= dtmin ends the step after one trial: coord = c0 + v0 dtmin with
    v0 the velocity at the start."""
    dln, mat = prm
    net = _glide_loop()
    n = int(net.numNode[0])
    ref = _glide_loop()
    O.deriv_(dln, mat, None, None, ref)
    v0 = ref.nodeVel[:, :n].copy()
    c0 = net.coord[:, :n].copy()
    ip = IntegrationParameters(dtmin=0.5, dtmax=1e15, abstol=10.0, reltol=1.0, maxiter=20)
    dt, time, step = O.integrate_(ip, IntegrationTime(dt=0.0), dln, mat, None, None, net)
    assert (dt, time, step) == (0.5, 0.5, 1)
    assert np.array_equal(net.coord[:, :n], c0 + v0 * 0.5)


def test_rejected_steps_halve_until_dtmin(prm):
    """abstol below any possible displacement rejects every trial: dt halves
    from dt0 until it is <= dtmin, and that trial's coordinates are kept."""
    dln, mat = prm
    net = _glide_loop()
    n = int(net.numNode[0])
    ref = _glide_loop()
    O.deriv_(dln, mat, None, None, ref)
    v0 = ref.nodeVel[:, :n].copy()
    c0 = net.coord[:, :n].copy()
    ip = IntegrationParameters(dtmin=1.0, dtmax=1e15, abstol=1e-300, reltol=1.0, maxiter=20)
    dt, _, _ = O.integrate_(ip, IntegrationTime(dt=100.0), dln, mat, None, None, net)
    want = 100.0
    while want > 1.0:
        want /= 2
    assert dt == want
    assert np.array_equal(net.coord[:, :n], c0 + v0 * want)


def test_glide_loop_shrinks_in_its_plane(prm):
    """Line tension shrinks a glide loop; the climb drag (1e10) keeps the
    motion in the glide plane."""
    dln, mat = prm
    net = _glide_loop()
    n = int(net.numNode[0])
    normal = net.slipPlane[:, 0] / np.linalg.norm(net.slipPlane[:, 0])
    c0 = net.coord[:, :n].copy()
    ctr = c0.mean(1, keepdims=True)
    r0 = np.linalg.norm(c0 - ctr, axis=0).mean()
    ip = IntegrationParameters(dtmin=1e-6, dtmax=1e15, abstol=10.0, reltol=1.0, maxiter=20)
    t = IntegrationTime(dt=100.0)
    for _ in range(3):
        dt, time, step = O.integrate_(ip, t, dln, mat, None, None, net)
        t = IntegrationTime(dt, time, step)
    d = net.coord[:, :n] - c0
    r1 = np.linalg.norm(net.coord[:, :n] - ctr, axis=0).mean()
    assert r1 < r0 - 1e-3
    assert np.abs(normal @ d).max() < 1e-6 * np.abs(d).max()
    assert t.step == 3 and t.time > 0


def test_fixed_nodes_do_not_move(prm):
    dln, mat = prm
    net = random_loop_network(4, seed=0xA3)
    n = int(net.numNode[0])
    lab = net.label.copy()
    lab[:n:5] = intFixDln
    net = DislocationNetwork._assemble(net.links, net.slipPlane, net.bVec, net.coord, lab, n,
                                       int(net.numSeg[0]), net.maxConnect)
    c0 = net.coord[:, :n].copy()
    ip = IntegrationParameters(dtmin=1e-6, dtmax=1e15, abstol=10.0, reltol=1.0, maxiter=5)
    O.integrate_(ip, IntegrationTime(dt=100.0), dln, mat, None, None, net)
    moved = np.abs(net.coord[:, :n] - c0).max(0) > 0
    assert not moved[::5].any() and moved.sum() > n // 2


def test_surface_mobile_nodes_are_refused(prm):
    dln, mat = prm
    net = _glide_loop()
    net.label[0] = 3
    with pytest.raises(ValueError, match="srfMobDln"):
        O.deriv_(dln, mat, None, None, net)
