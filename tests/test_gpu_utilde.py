"""GPU parity of calc_uTilde (ddd_u_tilde: DislocationFEM.jl:355-707) through
the C-ABI: against the reference's golden displacements at its tolerance and
against the CPU oracle on the same inputs.  The kernel evaluates the
oracle's expressions in the same order without contraction; acos, tan, atan
and log come from the device library (the oracle uses the host's), and
chunks of triangles are summed separately, so the bar is 1e-11 norm-wise
rather than bit identity."""
import numpy as np
import pytest

import pyoracle as O
import refnets as R
from dddhip import forces as G
from dddhip.generate import random_loop_network
from dddhip.types import DislocationNetwork, MaterialParameters, extDln, srfMobDln, srfFixDln
from metrics import normwise
from test_oracle_utilde import _mesh

pytestmark = pytest.mark.gpu

TOL = 1e-11


def _log(msg):
    print(msg, flush=True)


def test_golden_displacements():
    mat, m, bnd, fd, net = R.utilde_setup()
    u = G.calc_uTilde(m, bnd, mat, net)
    for c, key in enumerate(("utilde_uxTest", "utilde_uyTest", "utilde_uzTest")):
        assert normwise(u[c::3], R.g(key)) <= 5e-5
    assert normwise(u, O.calc_uTilde(m, bnd, mat, net)) <= TOL
    G.calc_uTilde_(fd, m, bnd, mat, net)
    assert normwise(fd.uTilde, R.g("utilde_uTildeTest")) <= 5e-5


def test_random_network_against_oracle():
    """128 loops x 32 segments in the 2000^3 cube, 1,500 points on the
    surface and inside."""
    mat = MaterialParameters(μ=1.0, ν=0.28)
    net = random_loop_network(128, seed=0x7A)
    m = _mesh((-300.0, -300.0, -300.0), (2300.0, 2300.0, 2300.0), mat)
    pts = np.random.default_rng(8).uniform(-300.0, 2300.0, size=(3, 1500))
    _log("uTilde 4096 segments x 1500 points")
    u = G.u_tilde_points(mat, net, m, pts)
    o = O.u_tilde_points(mat, net, m, pts)
    assert normwise(u, o) <= TOL
    per = np.linalg.norm(u - o, axis=0) / np.maximum(np.linalg.norm(o, axis=0), 1e-6 * np.abs(o).max())
    assert per.max() <= 1e-9


def test_external_and_surface_nodes_against_oracle():
    """External nodes (closure over projections onto the surface) and
    surface-external connections, through the same kernel."""
    mat = MaterialParameters(μ=1.0, ν=0.28)
    net = random_loop_network(16, seed=0x7B, ext_fraction=0.2)
    n = int(net.numNode[0])
    lab = net.label.copy()
    # surface nodes next to some external ones
    for i in range(int(net.numSeg[0])):
        a, b = net.links[:, i]
        if lab[a - 1] == extDln and lab[b - 1] not in (extDln,) and i % 2 == 0:
            lab[b - 1] = srfMobDln if i % 4 == 0 else srfFixDln
    net = DislocationNetwork._assemble(net.links, net.slipPlane, net.bVec, net.coord, lab, n,
                                       int(net.numSeg[0]), net.maxConnect)
    m = _mesh((0.0, 0.0, 0.0), (2000.0, 2000.0, 2000.0), mat)
    pts = np.random.default_rng(9).uniform(0.0, 2000.0, size=(3, 700))
    u = G.u_tilde_points(mat, net, m, pts)
    o = O.u_tilde_points(mat, net, m, pts)
    assert np.all(np.isfinite(o))
    assert normwise(u, o) <= TOL


def test_accumulate_and_empty():
    mat, m, bnd, fd, net = R.utilde_setup()
    pts = m.coord[:, bnd.uGammaDln - 1]
    u = G.u_tilde_points(mat, net, m, pts)
    assert G.u_tilde_points(mat, net, m, np.zeros((3, 0))).shape == (3, 0)
    empty = DislocationNetwork.from_arrays(np.zeros((2, 0), dtype=np.int64), np.zeros((3, 0)),
                                           np.zeros((3, 0)), [])
    assert np.all(G.u_tilde_points(mat, empty, m, pts) == 0)
