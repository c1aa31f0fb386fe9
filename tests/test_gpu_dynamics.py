"""GPU parity of the device-resident deriv!/integrate!
(ddd_net_upload_dynamics, ddd_net_deriv, ddd_net_integrate:
IntegrationProcessing.jl:13-168) through the C-ABI, against the CPU oracle's
restatement of the same functions.

deriv! is pinned to the reference's golden node forces and velocities and
to the oracle at identical coordinates.  For integrate! the reference holds
no trajectory, so the bar is agreement with the oracle's trajectory, whose
controller is pinned in closed form by test_oracle_integration.py; the two
force evaluations differ by summation order (~1e-15)."""
import numpy as np
import pytest

import pyoracle as O
import refnets as R
from dddhip import forces as G
from dddhip.generate import random_loop_network
from dddhip.types import (DislocationLoop, DislocationNetwork, IntegrationTime, SlipSystem,
                          extDln, intFixDln, intMobDln)
from metrics import isapprox, normwise, per_node_rel

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


def test_deriv_golden():
    """deriv! on the shear hexagon + prismatic pentagon (MobilityTest.jl
    network, PK = 0): node forces and mobile-node velocities at the golden
    values, fixed nodes (intFixDln) at rest."""
    dln, mat = R.params()
    net = R.mobility_network()
    G.deriv_(dln, mat, None, None, net)
    n = int(net.numNode[0])
    assert isapprox(net.nodeForce[:, :n].T, R.g("mobility_testForce"))
    fixed = np.asarray(net.label[:n]) == intFixDln
    assert fixed.any() and np.all(net.nodeVel[:, :n][:, fixed] == 0)
    gold = R.g("mobility_testVel").T
    assert isapprox(net.nodeVel[:, :n][:, ~fixed], gold[:, ~fixed], rtol=1e-5)
    ref = R.mobility_network()
    O.deriv_(dln, mat, None, None, ref)
    assert per_node_rel(net.segForce[:, :, :int(net.numSeg[0])],
                        ref.segForce[:, :, :int(ref.numSeg[0])])[0] <= 1e-10
    # same coordinates, so the drag tensors are identical; only the forces
    # differ by summation order
    assert per_node_rel(net.nodeVel[:, :n][:, ~fixed], ref.nodeVel[:, :n][:, ~fixed])[0] <= 1e-9


def _network():
    """20 loops x 32 segments (seed 0xA0, loops may intersect) with 5 %
    external nodes and every fifth remaining node fixed."""
    net = random_loop_network(20, seed=0xA0, ext_fraction=0.05)
    return _relabel(net, fixed_every=5)


def _relabel(net, fixed_every, ext=()):
    n = int(net.numNode[0])
    lab = net.label.copy()
    pick = np.zeros(len(lab), bool)
    pick[:n:fixed_every] = True
    lab[pick & (lab != extDln)] = intFixDln
    lab[list(ext)] = extDln
    return DislocationNetwork._assemble(net.links, net.slipPlane, net.bVec, net.coord, lab, n,
                                        int(net.numSeg[0]), net.maxConnect)


def _single_loop():
    """One 31-sided shear loop of radius 300 on slip system 4 with every
    seventh node fixed and nodes 1-2 external.  An odd polygon has no two
    parallel sides (the closest pair is pi/31 apart), which keeps every pair
    far from the reference's near-parallel switch: just above it the generic
    pair formula loses all its digits (a 1e-13 change of coordinates moved a
    pair's force 14-fold at 1 - c^2 = 2.6e-8), and networks of several loops
    reach that band as they move."""
    lp = DislocationLoop(loopType="shear", numSides=31, nodeSide=1, numLoops=1,
                         segLen=2 * 300.0 * np.sin(np.pi / 31) * np.ones(31), slipSystemIdx=4,
                         slipSystem=SlipSystem.bcc(), label=[intMobDln] * 31, buffer=0.0,
                         range=np.array([[0.0, 2000.0]] * 3), dist=np.full((3, 1), 0.5))
    return _relabel(DislocationNetwork.from_loops(lp, memBuffer=1), 7, ext=[0, 1])


def test_deriv_matches_oracle_on_intersecting_loops():
    """deriv! at identical coordinates on loops that cross each other."""
    dln, mat = R.params()
    net = _network()
    G.deriv_(dln, mat, None, None, net)
    ref = _network()
    O.deriv_(dln, mat, None, None, ref)
    n, ns = int(net.numNode[0]), int(net.numSeg[0])
    assert per_node_rel(net.segForce[:, :, :ns], ref.segForce[:, :, :ns])[0] <= 1e-10
    assert per_node_rel(net.nodeForce[:, :n], ref.nodeForce[:, :n])[0] <= 1e-10
    assert per_node_rel(net.nodeVel[:, :n], ref.nodeVel[:, :n])[0] <= 1e-9


# integrate! is sensitive to rounding by construction: the accepted step
# follows maxErr through the step factor, and velocities carry the drag
# tensor's conditioning (~1e10 with the sample's climb drag).  Measured on
# the oracle alone, a 1e-14 relative change of the start coordinates moves
# the displacement after 1-3 steps by ~1e-4 with the sample drag (climb 1e10,
# line 0) and by ~1e-6 with climb 1e3, line 1.  The device and the oracle
# differ by force summation order, a perturbation of that size, so the bars
# sit just above those spreads.
@pytest.mark.parametrize("drag,bar", [({"climb": 1e10, "line": 0.0}, 1e-3),
                                      ({"climb": 1e3, "line": 1.0}, 1e-5)],
                         ids=["sample-drag", "climb1e3-line1"])
def test_integrate_matches_oracle(drag, bar):
    dln, mat = R.params()
    dln.dragCoeffs = dict(dln.dragCoeffs, **drag)
    ip, t0 = R.integration()
    steps = 3
    net = _single_loop()
    n = int(net.numNode[0])
    c0 = net.coord[:, :n].copy()
    _log(f"integrate {drag} x {steps}: device")
    r = G.ResidentNetwork(dln, mat, net)
    r.upload_dynamics(dln, net)
    t = t0
    for _ in range(steps):
        t = r.integrate(ip, t)
    cg, vg, fg = r.download_nodes()
    _log(f"integrate {drag} x {steps}: oracle")
    ref = _single_loop()
    to = t0
    for _ in range(steps):
        to = IntegrationTime(*O.integrate_(ip, to, dln, mat, None, None, ref))
    dg, do = cg - c0, ref.coord[:, :n] - c0
    err = np.abs(dg - do).max() / np.abs(do).max()
    _log(f"  device t={t} oracle t={to} displacement err {err:.3e} "
         f"max displacement {np.abs(do).max():.3e}")
    assert t.step == to.step == t0.step + steps
    assert abs(t.dt - to.dt) <= bar * to.dt and abs(t.time - to.time) <= bar * to.time
    assert err <= bar
    assert per_node_rel(fg, ref.nodeForce[:, :n])[1] <= bar
    lab = np.asarray(net.label[:n])
    still = (lab == intFixDln) | (lab == extDln)
    assert np.all(dg[:, still] == 0) and np.all(vg[:, still] == 0)
    assert np.abs(dg[:, ~still]).max(0).min() > 0


def test_drop_in_integrate_equals_resident():
    """integrate_ (upload, one step, download into the network) gives the
    resident path's bits, and fills segForce, nodeForce and nodeVel."""
    dln, mat = R.params()
    ip, t0 = R.integration()
    a = _network()
    t1 = G.integrate_(ip, t0, dln, mat, None, None, a)
    b = _network()
    r = G.ResidentNetwork(dln, mat, b)
    r.upload_dynamics(dln, b)
    t2 = r.integrate(ip, t0)
    cb, vb, fb = r.download_nodes()
    n = int(a.numNode[0])
    assert (t1.dt, t1.time, t1.step) == (t2.dt, t2.time, t2.step)
    assert np.array_equal(a.coord[:, :n], cb) and np.array_equal(a.nodeVel[:, :n], vb)
    assert np.array_equal(a.nodeForce[:, :n], fb)
    assert np.array_equal(a.segForce[:, :, :int(a.numSeg[0])], r.download())


def test_deriv_in_mixed_mode():
    """The resident dynamics run in the mixed-precision mode too: velocities
    within the mixed path's force error of the FP64 ones."""
    dln, mat = R.params()
    net = random_loop_network(64, seed=0xC1)
    n = int(net.numNode[0])
    r = G.ResidentNetwork(dln, mat, net)
    r.upload_dynamics(dln, net)
    r.deriv()
    _, v64, f64 = r.download_nodes()
    r.set_mode("mixed")
    try:
        r.deriv()
        _, vmx, fmx = r.download_nodes()
    finally:
        r.set_mode("fp64")   # the mode belongs to the shared context
    assert per_node_rel(fmx, f64)[1] <= 1e-5
    assert per_node_rel(vmx, v64)[1] <= 1e-4


def test_dynamics_errors():
    dln, mat = R.params()
    net = _network()
    r = G.ResidentNetwork(dln, mat, net)
    with pytest.raises(RuntimeError, match="no dynamics state"):
        r.deriv()
    bad = _network()
    bad.label[3] = 3   # srfMobDln
    r2 = G.ResidentNetwork(dln, mat, bad)
    with pytest.raises(RuntimeError, match="srfMobDln"):
        r2.upload_dynamics(dln, bad)
