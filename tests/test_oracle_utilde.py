"""CPU tests of the oracle's calc_uTilde (DislocationFEM.jl:355-707).

Pinned by the reference's golden displacements (DislocationFEMTest.jl:
270-1141, rtol 5e-5 there; the oracle agrees to ~1e-12) for internal
segments, and by what the mathematics of a dislocation fixes without the
oracle's formula: the displacement of a closed loop depends on the closure
point only through whole Burgers vectors (the solid angle jumps by 4 pi when
a closure cone sweeps over a point; the line terms of the closure edges
cancel between neighbouring triangles), and it jumps by exactly b across the
cut surface.  The closure of external segments (projection onto the mesh
surface, :437-528) has no reference vectors: parity unpinned beyond the
consistency checks here and the device comparison."""
import numpy as np
import pytest

import pyoracle as O
import refnets as R
from dddhip.types import (DislocationLoop, DislocationNetwork, MaterialParameters,
                          RegularCuboidMesh, SlipSystem, intMobDln, extDln, srfMobDln)
from metrics import normwise


def test_golden_displacements():
    mat, m, bnd, fd, net = R.utilde_setup()
    u = O.calc_uTilde(m, bnd, mat, net)
    for c, key in enumerate(("utilde_uxTest", "utilde_uyTest", "utilde_uzTest")):
        assert normwise(u[c::3], R.g(key)) <= 1e-10
    O.calc_uTilde_(fd, m, bnd, mat, net)
    assert normwise(fd.uTilde, R.g("utilde_uTildeTest")) <= 1e-10
    # the in-place form scatters value k into dof uDofsDln[k] (x dofs, then
    # y, then z), so only the first third lands on the node's own x dof
    assert np.array_equal(fd.uTilde[np.asarray(bnd.uDofsDln) - 1], u)


def _loop(kind, center, sides=8, R0=300.0, slip=None):
    ss = slip or SlipSystem(np.array([[-1.0], [1.0], [0.0]]), np.array([[1.0], [1.0], [1.0]]))
    return DislocationLoop(loopType=kind, numSides=sides, nodeSide=1, numLoops=1,
                           segLen=2 * R0 * np.sin(np.pi / sides) * np.ones(sides),
                           slipSystemIdx=1, slipSystem=ss, label=[intMobDln] * sides, buffer=0.0,
                           range=np.array([[c, c] for c in center]), dist="zeros")


def _mesh(lo, hi, mat):
    m = RegularCuboidMesh(dx=hi[0] - lo[0], dy=hi[1] - lo[1], dz=hi[2] - lo[2], mx=2, my=2,
                          mz=2, μ=mat.μ, ν=mat.ν)
    shift = np.asarray(lo, dtype=float)[:, None]
    m.vertices = np.asfortranarray(m.vertices + shift)
    m.faceMidPt = np.asfortranarray(m.faceMidPt + shift)
    return m


def test_closure_point_changes_u_by_whole_burgers_vectors():
    mat = MaterialParameters(μ=1.0, ν=0.28)
    net = DislocationNetwork.from_loops([_loop("shear", (1000.0, 1000.0, 1500.0)),
                                         _loop("prism", (1100.0, 900.0, 1400.0), sides=7)])
    rng = np.random.default_rng(5)
    pts = rng.uniform(0.0, 2500.0, size=(3, 400))
    ua = O.u_tilde_points(mat, net, _mesh((0, 0, 0), (2000, 2000, 3000), mat), pts)
    ub = O.u_tilde_points(mat, net, _mesh((-700, 300, 100), (2100, 2500, 2600), mat), pts)
    b = np.array([1.0, 1.0, 1.0]) / np.sqrt(3)     # both loops carry this b
    d = ua - ub
    k = np.round(d.T @ b)                            # whole Burgers vectors
    assert np.abs(d - np.outer(b, k)).max() <= 1e-9
    assert np.any(k != 0) and np.any(k == 0)         # some points saw a cone sweep
    assert np.abs(ua).max() > 1e-3


@pytest.mark.parametrize("kind", ["shear", "prism"])
def test_jump_of_b_across_the_cut(kind):
    """Closure at the loop centre: the cut is the loop's flat polygon, and
    u(above) - u(below) = b at points inside it; outside the loop u is
    continuous through the plane."""
    mat = MaterialParameters(μ=1.0, ν=0.28)
    ctr = np.array([1000.0, 1000.0, 1500.0])
    net = DislocationNetwork.from_loops([_loop(kind, ctr, sides=8)])
    m = _mesh(ctr - 1000.0, ctr + 1000.0, mat)       # centroid = loop centre
    n = net.slipPlane[:, 0] / np.linalg.norm(net.slipPlane[:, 0])
    if kind == "prism":
        n = net.bVec[:, 0] / np.linalg.norm(net.bVec[:, 0])
    e = np.cross(n, [0.3, 0.5, 0.7])
    e /= np.linalg.norm(e)
    b = net.bVec[:, 0]
    for r, inside in ((37.0, True), (120.0, True), (700.0, False)):
        base = ctr + r * e
        up = O.u_tilde_points(mat, net, m, (base + 1e-6 * n)[:, None])[:, 0]
        dn = O.u_tilde_points(mat, net, m, (base - 1e-6 * n)[:, None])[:, 0]
        jump = up - dn
        if inside:
            assert np.allclose(np.abs(jump @ b), b @ b, rtol=0, atol=1e-6)
            assert np.linalg.norm(jump - (jump @ b) / (b @ b) * b) <= 1e-6
        else:
            assert np.linalg.norm(jump) <= 1e-6


def test_external_segments_run_and_change_u():
    """A loop with external nodes and a surface node: the closure over the
    projected points runs, stays finite and differs from the all-internal
    result (parity unpinned: no reference vectors exist for it)."""
    mat = MaterialParameters(μ=1.0, ν=0.28)
    net = DislocationNetwork.from_loops([_loop("shear", (1000.0, 1000.0, 1500.0), sides=8)])
    m = _mesh((0, 0, 0), (2000, 2000, 3000), mat)
    pts = np.random.default_rng(6).uniform(0.0, 2000.0, size=(3, 50))
    u0 = O.u_tilde_points(mat, net, m, pts)
    lab = net.label.copy()
    lab[0:2] = extDln
    lab[2] = srfMobDln
    n = int(net.numNode[0])
    ext = DislocationNetwork._assemble(net.links, net.slipPlane, net.bVec, net.coord, lab, n,
                                       int(net.numSeg[0]), net.maxConnect)
    u1 = O.u_tilde_points(mat, ext, m, pts)
    assert np.all(np.isfinite(u1)) and not np.allclose(u0, u1)
