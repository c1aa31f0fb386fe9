"""Seeded synthetic networks for the force-path workloads.

Loops are built exactly as the reference builds them (regular polygonal shear
loops, DislocationConstructor.jl:196-298, placed by makeNetwork!,
:645-688), with the reference's ``rand`` replaced by a counter-based SplitMix64
stream so that every machine generates bit-identical networks from a seed.

Draw order per loop k (five uniforms u[5k .. 5k+4] in [0, 1)):
  slip system s = 1 + floor(12 u0) of the 12 BCC {110}<111> systems,
  radius R = 100 + 200 u1, polygon side 2 R sin(pi / numSides),
  centre (2000 u2, 2000 u3, 2000 u4) in the 2000^3 cube.
"""
from __future__ import annotations

import math

import numpy as np

from .types import (DislocationLoop, DislocationNetwork, SlipSystem, intMobDln, extDln,
                    RegularCuboidMesh, ForceDisplacement)

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64_uniform(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """u_k = (mix(seed + (offset + k + 1) * gamma) >> 11) * 2^-53, k = 0..n-1."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def random_loop_network(numLoops: int, numSides: int = 32, seed: int = 0xA0,
                        side: float = 2000.0, rmin: float = 100.0, rmax: float = 300.0,
                        memBuffer=1, ext_fraction: float = 0.0) -> DislocationNetwork:
    """``numLoops`` shear loops of ``numSides`` sides on random BCC slip
    systems, radius U(rmin, rmax), centres uniform in [0, side]^3.

    ``ext_fraction`` > 0 relabels that share of nodes as external (a test of
    the extSeg handling); the reference generator never does."""
    u = splitmix64_uniform(seed, 5 * numLoops)
    ss = SlipSystem.bcc()
    loops = []
    for k in range(numLoops):
        s = 1 + min(11, int(math.floor(12 * u[5 * k])))
        R = rmin + (rmax - rmin) * u[5 * k + 1]
        L = 2 * R * math.sin(math.pi / numSides)
        disp = np.array([[u[5 * k + 2]], [u[5 * k + 3]], [u[5 * k + 4]]])
        loops.append(DislocationLoop(
            loopType="shear", numSides=numSides, nodeSide=1, numLoops=1,
            segLen=L * np.ones(numSides), slipSystemIdx=s, slipSystem=ss,
            label=[intMobDln] * numSides, buffer=0.0,
            range=np.array([[0.0, side], [0.0, side], [0.0, side]]), dist=disp))
    net = DislocationNetwork.from_loops(loops, memBuffer=memBuffer)
    if ext_fraction > 0:
        n = int(net.numNode[0])
        v = splitmix64_uniform(seed ^ 0x5EED, n)
        lab = net.label.copy()
        lab[:n][v < ext_fraction] = extDln
        net = DislocationNetwork._assemble(net.links, net.slipPlane, net.bVec, net.coord, lab,
                                           int(net.numNode[0]), int(net.numSeg[0]),
                                           net.maxConnect)
    return net


def face_grid_points(n_per_edge: int = 128, side: float = 2000.0) -> np.ndarray:
    """Cell-centred n x n grids on the 6 faces of [0, side]^3: 6 n^2 points,
    3 x P column-major, faces in the order x=0, x=side, y=0, y=side, z=0,
    z=side; on each face the first free coordinate runs fastest."""
    h = side / n_per_edge
    g = (np.arange(n_per_edge) + 0.5) * h
    A, B = np.meshgrid(g, g, indexing="xy")  # A fastest
    a, b = A.ravel(), B.ravel()
    faces = []
    for axis in range(3):
        others = [i for i in range(3) if i != axis]
        for val in (0.0, side):
            p = np.zeros((3, a.size))
            p[axis] = val
            p[others[0]] = a
            p[others[1]] = b
            faces.append(p)
    return np.asfortranarray(np.concatenate(faces, axis=1))


def random_uhat(mesh: RegularCuboidMesh, seed: int = 0xF0, density: float = 0.05,
                scale: float = 1000.0) -> ForceDisplacement:
    """A seeded sparse FE correction displacement (dense storage), like the
    sparse uHat vectors of SegmentForcesTest.jl:65-81."""
    n = 3 * mesh.numNode
    u = splitmix64_uniform(seed, 2 * n)
    uHat = np.where(u[:n] < density, u[n:] * scale, 0.0)
    return ForceDisplacement(uHat=uHat)


def clustered_loop_network(numLoops: int = 8192, numSides: int = 32, seed: int = 0xE0,
                           clusters: int = 8, cluster_frac: float = 0.7, sigma: float = 150.0,
                           side: float = 2000.0, rmin: float = 100.0,
                           rmax: float = 300.0) -> DislocationNetwork:
    """A clustered network: ``cluster_frac`` of the loops form ``clusters``
    pile-ups, the rest are placed as by random_loop_network.  The loops of a
    pile-up share one BCC slip system (one Burgers vector and glide plane,
    so their sides are parallel) and have centres drawn from N(c, sigma^2 I)
    around a pile-up centre c uniform in [side/4, 3 side/4]^3.  Radius
    U(rmin, rmax) for every loop.

    Draws: 4 uniforms per pile-up (slip system, centre), then 6 per loop
    (radius, three for the uniform centre or two Box-Muller pairs' worth for
    the Gaussian one, slip system for uniform loops)."""
    ss = SlipSystem.bcc()
    nc = int(round(cluster_frac * numLoops))
    uc = splitmix64_uniform(seed, 4 * clusters)
    cl_sys = [1 + min(11, int(math.floor(12 * uc[4 * c]))) for c in range(clusters)]
    cl_ctr = [side * (0.25 + 0.5 * uc[4 * c + 1:4 * c + 4]) for c in range(clusters)]
    u = splitmix64_uniform(seed, 6 * numLoops, offset=4 * clusters)
    loops = []
    for k in range(numLoops):
        v = u[6 * k:6 * k + 6]
        R = rmin + (rmax - rmin) * v[0]
        if k < nc:
            c = k % clusters
            s = cl_sys[c]
            # Box-Muller: three normals from two uniform pairs
            r1 = math.sqrt(-2.0 * math.log(max(v[1], 1e-300)))
            r2 = math.sqrt(-2.0 * math.log(max(v[3], 1e-300)))
            g = np.array([r1 * math.cos(2 * math.pi * v[2]), r1 * math.sin(2 * math.pi * v[2]),
                          r2 * math.cos(2 * math.pi * v[4])])
            centre = cl_ctr[c] + sigma * g
        else:
            s = 1 + min(11, int(math.floor(12 * v[5])))
            centre = side * v[1:4]
        L = 2 * R * math.sin(math.pi / numSides)
        loops.append(DislocationLoop(
            loopType="shear", numSides=numSides, nodeSide=1, numLoops=1,
            segLen=L * np.ones(numSides), slipSystemIdx=s, slipSystem=ss,
            label=[intMobDln] * numSides, buffer=0.0,
            range=np.array([[0.0, 1.0], [0.0, 1.0], [0.0, 1.0]]),
            dist=centre.reshape(3, 1)))
    return DislocationNetwork.from_loops(loops)


def surface_points(n: int, seed: int = 0xE5, side: float = 2000.0) -> np.ndarray:
    """n points uniform on the surface of [0, side]^3 (face chosen uniformly),
    3 x n column-major."""
    u = splitmix64_uniform(seed, 3 * n).reshape(n, 3)
    face = np.minimum(5, (6 * u[:, 0]).astype(np.int64))
    axis, hi = face // 2, face % 2
    p = np.zeros((n, 3))
    other = np.array([[1, 2], [0, 2], [0, 1]])[axis]
    p[np.arange(n), axis] = hi * side
    p[np.arange(n), other[:, 0]] = side * u[:, 1]
    p[np.arange(n), other[:, 1]] = side * u[:, 2]
    return np.asfortranarray(p.T)
