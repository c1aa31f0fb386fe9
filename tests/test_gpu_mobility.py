"""GPU parity of the BCC mobility (ddd_mobility_bcc: dlnMobility /
dlnMobility!, DislocationMobility.jl:16-341) through the C-ABI.

Bar: the reference's golden velocities at its own tolerance (rtol 1e-5,
MobilityTest.jl:92), and bit-identical node forces and velocities against
the CPU oracle on the same segment forces: the kernel forms the drag tensor
and its Cramer solve with the oracle's operations in the oracle's order and
without contraction, which matters because the tensor's condition number
(up to ~1e10 with climb drag 1e10) turns any rounding difference into
~1e-6 of velocity."""
import numpy as np
import pytest

import pyoracle as O
import refnets as R
from dddhip import forces as G
from dddhip.generate import random_loop_network
from dddhip.types import DislocationParameters, MaterialParameters
from metrics import isapprox

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


def _mixed_branch_network(numLoops, seed):
    """Random shear loops whose links are re-assigned Burgers vectors from
    every branch of the law: <111> unit (screw family), (2,1,1) and (3,2,1)
    families, non-unit <111> (|b| < 1 and |b| > 1), vectors with a zero
    component (edge law); plus collapsed links (zero length)."""
    net = random_loop_network(numLoops, seed=seed)
    rng = np.random.default_rng(seed)
    nl = int(net.numSeg[0])
    choices = [np.array([1, 1, 1]) / np.sqrt(3), np.array([-1, 1, 1]) / np.sqrt(3),
               np.array([2, -1, 1]) / np.sqrt(6), np.array([3, 2, -1]) / np.sqrt(14),
               np.array([0.5, 0.5, 0.5]), np.array([1.0, 1.0, 1.0]),
               np.array([1.0, 0.0, 0.0]), np.array([1, 1, 0]) / np.sqrt(2)]
    pick = rng.integers(0, len(choices) + 2, size=nl)
    for i in range(nl):
        if pick[i] < len(choices):
            net.bVec[:, i] = choices[pick[i]]
    # collapse a few links: node 2 of the link onto node 1
    for i in rng.choice(nl, size=max(1, nl // 97), replace=False):
        n1, n2 = net.links[:, i]
        net.coord[:, n2 - 1] = net.coord[:, n1 - 1]
    return net


def _random_segforce(net, seed):
    rng = np.random.default_rng(seed)
    sf = np.zeros_like(net.segForce)
    n = int(net.numSeg[0])
    sf[:, :, :n] = rng.standard_normal((3, 2, n))
    return sf


def _dln(drag):
    return DislocationParameters(coreRad=90.78, coreEnergy=1.0, dragCoeffs=drag)


DRAGS = [{"edge": 1.0, "screw": 2.0, "climb": 1e10, "line": 0.0},
         {"edge": 1.0, "screw": 2.0, "climb": 1e3, "line": 0.5}]


def test_golden_velocity():
    """calcSegForce! then dlnMobility on the device (MobilityTest.jl:51-98)."""
    dln, mat = R.params()
    net = R.mobility_network()
    m = R.mesh()
    fd = R.ForceDisplacement(np.zeros(3 * m.numNode))
    G.calcSegForce_(dln, mat, m, fd, net)
    force, vel = G.dlnMobility(dln, mat, net)
    assert isapprox(vel.T, R.g("mobility_testVel"), rtol=1e-5)
    assert isapprox(force.T, R.g("mobility_testForce"))
    of, ov = O.dlnMobility(dln, mat, net)   # same segment forces on the CPU
    assert np.array_equal(force, of) and np.array_equal(vel, ov)
    # in place (MobilityTest.jl:60-61,93-94)
    net2 = R.mobility_network()
    net2.segForce[:] = net.segForce
    G.dlnMobility_(dln, mat, net2)
    n = int(net.numNode[0])
    assert np.array_equal(net2.nodeForce[:, :n], force)
    assert np.array_equal(net2.nodeVel[:, :n], vel)
    # one node through idx (MobilityTest.jl:95-98)
    for node in (1, 5, n):
        fi, vi = G.dlnMobility(dln, mat, net, node)
        assert fi.shape == (3, 1)
        assert np.array_equal(fi[:, 0], force[:, node - 1])
        assert np.array_equal(vi[:, 0], vel[:, node - 1])


@pytest.mark.parametrize("drag", DRAGS, ids=["climb1e10", "climb1e3"])
@pytest.mark.parametrize("sigmaPN", [0.0, 0.05])
def test_mixed_branches_bit_exact(drag, sigmaPN):
    net = _mixed_branch_network(64, seed=0x3A)
    net.segForce[:] = _random_segforce(net, 7)
    dln = _dln(drag)
    mat = MaterialParameters(μ=1.0, ν=0.28, σPN=sigmaPN)
    _log(f"mobility mixed branches drag={drag} sigmaPN={sigmaPN}")
    f, v = G.dlnMobility(dln, mat, net)
    of, ov = O.dlnMobility(dln, mat, net)
    assert np.array_equal(f, of)
    assert np.array_equal(np.isnan(v), np.isnan(ov))
    assert np.array_equal(np.nan_to_num(v), np.nan_to_num(ov))
    if sigmaPN > 0:   # the cut removed some connection forces
        f0, _ = O.dlnMobility(dln, MaterialParameters(μ=1.0, ν=0.28), net)
        assert not np.array_equal(f, f0)


def test_idx_and_in_place_columns():
    """A node list: out-of-place fills columns 1..length(idx); in place writes
    columns idx[k] and leaves the others untouched (DislocationMobility.jl:
    40-46,164-165 and :209-215,336-337)."""
    net = _mixed_branch_network(16, seed=0x3B)
    net.segForce[:] = _random_segforce(net, 8)
    dln = _dln(DRAGS[1])
    mat = MaterialParameters(μ=1.0, ν=0.28)
    n = int(net.numNode[0])
    idx = np.array([n, 3, 17, 1, 100, 3], dtype=np.int64)
    f, v = G.dlnMobility(dln, mat, net, idx)
    of, ov = O.dlnMobility(dln, mat, net, idx=idx)
    assert f.shape == (3, len(idx))
    assert np.array_equal(f, of) and np.array_equal(v, ov)
    net.nodeForce[:] = -7.0
    net.nodeVel[:] = -9.0
    G.dlnMobility_(dln, mat, net, range(5, 40))
    fa, va = G.dlnMobility(dln, mat, net)
    assert np.array_equal(net.nodeForce[:, 4:39], fa[:, 4:39])
    assert np.array_equal(net.nodeVel[:, 4:39], va[:, 4:39])
    assert np.all(net.nodeForce[:, :4] == -7.0) and np.all(net.nodeForce[:, 39:] == -7.0)
    assert np.all(net.nodeVel[:, :4] == -9.0) and np.all(net.nodeVel[:, 39:] == -9.0)


def test_full_size_network_bit_exact():
    """131,072 nodes (4,096 loops x 32), random segment forces."""
    dln, mat = R.params()
    net = random_loop_network(4096, seed=0xD0)
    net.segForce[:] = _random_segforce(net, 9)
    _log("mobility 131072 nodes")
    f, v = G.dlnMobility(dln, mat, net)
    of, ov = O.dlnMobility(dln, mat, net)
    assert np.array_equal(f, of) and np.array_equal(v, ov)


def test_errors_are_reported():
    dln, mat = R.params()
    net = R.mobility_network()
    with pytest.raises(RuntimeError, match="idx outside"):
        G.dlnMobility(dln, mat, net, [0])
    with pytest.raises(RuntimeError, match="idx outside"):
        G.dlnMobility(dln, mat, net, [int(net.numNode[0]) + 1])
    bad = DislocationParameters(coreRad=90.78, dragCoeffs={"edge": 1.0, "screw": 2.0,
                                                           "climb": 1e9})
    with pytest.raises(KeyError):
        G.dlnMobility(bad, mat, net)
