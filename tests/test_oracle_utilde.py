"""CPU tests of the oracle's calc_uTilde (DislocationFEM.jl:355-707).

Pinned by the reference's golden displacements (DislocationFEMTest.jl:
270-1141, rtol 5e-5 there; the oracle agrees to ~1e-12) for internal
segments, and by what the mathematics of a dislocation fixes without the
oracle's formula: the displacement of a closed loop depends on the closure
point only through whole Burgers vectors (the solid angle jumps by 4 pi when
a closure cone sweeps over a point; the line terms of the closure edges
cancel between neighbouring triangles), and it jumps by exactly b across the
cut surface.  The closure of external segments (projection onto the mesh
surface, :437-528) has no reference vectors; it is pinned by geometry: the
surface points in closed form (one by a hand-worked trace of the
reference's walk) and the closure polygon's solid angle by an independent
triangulation and formula, equal up to whole Burgers vectors."""
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
    result (its geometry is pinned case by case below)."""
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


# ---- the closure of segments with external nodes (:437-528) ----------------
# For an external segment A -> B the reference finds A', B', the points of the
# mesh surface its ends were projected out from (the nearest intersection of
# the lines through A along the normals of faces 1, 3, 5 -- -y, +y, -z -- with
# the volume; a surface node linked to an external node by a virtual segment
# takes that role; with no intersection, a walk over the face planes), then
# adds the solid-angle terms of the triangle (A', B', C) and of the closed
# polygon A -> B -> B' -> A' fanned from C' = (A + B') / 2.  Everything but
# the points is geometry: the solid angle of a closed polygon does not depend
# on its triangulation except by whole 4 pi, so the displacement must equal
# b / 4 pi (Omega(A', B', C) + Omega(A, B, B') + Omega(A, B', A')) up to whole
# Burgers vectors, with Omega from the Van Oosterom-Strackee formula (not the
# oracle's l'Huilier evaluation) and A', B' fixed in closed form or by hand.

def _vos(A, B, Cc, p):
    """Solid angle of triangle (A, B, C) seen from points p (3 x P), signed by
    the orientation, Van Oosterom & Strackee, IEEE TBME 30 (1983) 125."""
    R1, R2, R3 = (np.asarray(X, float)[:, None] - p for X in (A, B, Cc))
    n1, n2, n3 = (np.linalg.norm(R, axis=0) for R in (R1, R2, R3))
    num = np.einsum("ij,ij->j", R1, np.cross(R2.T, R3.T).T)
    den = (n1 * n2 * n3 + np.einsum("ij,ij->j", R1, R2) * n3 + np.einsum("ij,ij->j", R1, R3) * n2
           + np.einsum("ij,ij->j", R2, R3) * n1)
    return 2.0 * np.arctan2(num, den)


def _closure_u(A, B, Ap, Bp, Cc, b, p):
    om = _vos(Ap, Bp, Cc, p) + _vos(A, B, Bp, p) + _vos(A, Bp, Ap, p)
    return np.outer(b, om) / (4 * np.pi)


def _ext_network(nodes, labels, links, b):
    from dddhip.types import DislocationNetwork
    coord = np.asarray(nodes, dtype=float).T
    L = np.asarray(links, dtype=np.int64).T
    bv = np.tile(np.asarray(b, float)[:, None], (1, L.shape[1]))
    return DislocationNetwork.from_arrays(L, bv, coord, labels)


# mesh [0, 2000] x [0, 2000] x [0, 3000], centroid (1000, 1000, 1500)
_BOX_C = np.array([1000.0, 1000.0, 1500.0])
_B = np.array([1.0, -1.0, 1.0]) / np.sqrt(3)

EXTERNAL_CASES = {
    # both ends above the top face: projected straight down onto z = 3000
    "above_top": ([(700, 800, 3300), (1300, 900, 3400)], [extDln, extDln], [(1, 2)], 0,
                  (700, 800, 3000), (1300, 900, 3000)),
    # an internal end 100 below the top: its nearest face along the y and z
    # lines is z = 3000 (y faces 800 and 1200 away)
    "internal_end": ([(700, 800, 2900), (1300, 900, 3400)], [intMobDln, extDln], [(1, 2)], 0,
                     (700, 800, 3000), (1300, 900, 3000)),
    # beyond the y = 0 face: the y line meets the volume at y = 0
    "beyond_y0": ([(900, -150, 1200), (1100, -250, 1400)], [extDln, extDln], [(1, 2)], 0,
                  (900, 0, 1200), (1100, 0, 1400)),
    # the external end 3 is linked to the surface node 1 by a virtual segment
    # (listed first, so its connection is recorded before segment 3 -> 2 is
    # evaluated); 3 then takes node 1 as its surface point
    "surface_node": ([(1000, 1000, 3000), (1400, 1100, 3300), (1000, 1000, 3200)],
                     [srfMobDln, extDln, extDln], [(1, 3), (3, 2)], 1,
                     (1000, 1000, 3000), (1400, 1100, 3000)),
    # beyond the x = 2000 face: faces 1, 3, 5 are searched (the y and z lines
    # stay outside the volume), so the reference walks over the face planes,
    # keeping the candidate closest to the centroid; worked by hand from
    # :470-489 for A = (2100, 1000, 1500): face 1 gives (2100, 2000, 1500)
    # (1,486.6 from C, the first candidate), faces 2-6 give (-2000, 2000,
    # 1500), (2100, -2000, 1500), (4200, 2000, 1500), (2100, 2000, 3000),
    # (2100, 2000, -3000), all farther; B lies above the top face
    "corner_walk": ([(2100, 1000, 1500), (1500, 1200, 3200)], [extDln, extDln], [(1, 2)], 0,
                    (2100, 2000, 1500), (1500, 1200, 3000)),
}


@pytest.mark.parametrize("case", sorted(EXTERNAL_CASES))
def test_external_closure_against_geometry(case):
    nodes, labels, links, k, Ap, Bp = EXTERNAL_CASES[case]
    mat = MaterialParameters(μ=1.0, ν=0.28)
    net = _ext_network(nodes, labels, links, _B)
    m = _mesh((0, 0, 0), (2000, 2000, 3000), mat)
    pts = np.random.default_rng(8).uniform((0, 0, 0), (2000, 2000, 3000), size=(400, 3)).T
    pts = np.asfortranarray(pts)
    u = O.u_tilde_points(mat, net, m, pts)
    A = np.asarray(nodes[links[k][0] - 1], float)
    B = np.asarray(nodes[links[k][1] - 1], float)
    want = _closure_u(A, B, np.asarray(Ap, float), np.asarray(Bp, float), _BOX_C, _B, pts)
    d = u - want
    kk = np.round(d.T @ _B)
    assert np.abs(d - np.outer(_B, kk)).max() <= 1e-10
    assert np.abs(want).max() > 1e-3
    # a wrong surface point (here: the other end's) is not a whole-b shift
    bad = _closure_u(A, B, np.asarray(Bp, float), np.asarray(Bp, float), _BOX_C, _B, pts)
    db = u - bad
    assert np.abs(db - np.outer(_B, np.round(db.T @ _B))).max() > 1e-4
