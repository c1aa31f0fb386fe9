"""Host-side mirror of the reference's parameter and network structs.

The arrays keep the reference's Julia memory layout (column-major, 1-based
Int64 indices) so that the C-ABI sees exactly what a Julia ``ccall`` would
pass.  Only the fields the force path reads are mirrored.

Reference: src/Type/DislocationType.jl:24-32,205-255,278-340,
src/Type/MaterialType.jl:37-53, src/Type/MaterialConstructor.jl:13-48,
src/Type/DislocationConstructor.jl:55-105,196-298,477-551,645-688,
src/Type/DislocationUtil.jl:96-163,193-295, src/Misc/Misc.jl:143-207.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

# nodeTypeDln (DislocationType.jl:24-32)
noneDln, intMobDln, intFixDln, srfMobDln, srfFixDln, extDln, tmpDln = range(7)


@dataclass
class MaterialParameters:
    """MaterialParameters (MaterialType.jl:37-53) with the derived constants of
    MaterialConstructor.jl:20-29."""

    μ: float = 1.0
    ν: float = 0.2
    μMag: float = 1.0
    σPN: float = 0.0

    def __post_init__(self):
        μ, ν = float(self.μ), float(self.ν)
        self.omνInv = 1 / (1 - ν)
        self.opνInv = 1 / (1 + ν)
        self.νomνInv = ν * self.omνInv
        self.νopνInv = ν * self.opνInv
        self.μ4π = μ / (4 * math.pi)
        self.μ8π = self.μ4π / 2
        self.μ4πν = self.μ4π * self.omνInv
        self.νμ4πν = ν * self.μ4πν
        self.omνInv8π = self.omνInv / (8 * math.pi)
        self.om2νomνInv8π = (1 - 2 * ν) * self.omνInv8π


@dataclass
class DislocationParameters:
    """The fields of DislocationParameters (DislocationType.jl:205-255) that the
    force path reads.  ``coreRadSq`` is derived as in
    DislocationConstructor.jl:84.  ``parGPU`` is the drop-in switch the Julia
    shim routes on."""

    coreRad: float = 1.0
    coreEnergy: Optional[float] = None
    parCPU: bool = False
    parGPU: bool = True
    # mobility law and drag coefficients (DislocationType.jl:151-152); the
    # keyword default of DislocationConstructor.jl:57 has no "line" entry, so
    # the mobility function fails on it unless the caller supplies one, as
    # in the reference (DislocationMobility.jl:28).
    mobility: str = "mobBCC"
    dragCoeffs: dict = field(default_factory=lambda: {"edge": 1.0, "screw": 2.0, "climb": 1e9})
    # collision distance; keyword default minSegLen / 2 = coreRad
    # (DislocationConstructor.jl:57-67), squared as at :95
    collisionDist: Optional[float] = None

    def __post_init__(self):
        self.coreRad = float(self.coreRad)
        self.coreRadSq = self.coreRad ** 2
        if self.collisionDist is None:
            self.collisionDist = 2 * self.coreRad / 2
        self.collisionDistSq = float(self.collisionDist) ** 2
        if self.coreEnergy is None:
            # keyword default, DislocationConstructor.jl:60
            self.coreEnergy = 1 / (4 * math.pi) * math.log(self.coreRad / 0.1)


@dataclass
class IntegrationParameters:
    """IntegrationParameters (IntegrationType.jl:28-39) with the keyword
    defaults of IntegrationConstructor.jl:17-42; the only method is
    AdaptiveEulerTrapezoid."""

    method: str = "AdaptiveEulerTrapezoid"
    tmin: float = 0.0
    tmax: float = 1e13
    dtmin: float = 1e-3
    dtmax: float = math.inf
    abstol: float = 1e-6
    reltol: float = 1e-6
    maxchange: float = 1.2
    exponent: float = 20.0
    maxiter: int = 10


@dataclass
class IntegrationTime:
    """IntegrationTime (IntegrationType.jl:51-55)."""

    dt: float = 0.0
    time: float = 0.0
    step: int = 0


# data/slipSystems/BCC.json: 12 {110}<111> slip systems (plane normal, Burgers
# vector), in the file's order.  Kept as a table because the generator needs it
# where the reference tree is absent.
BCC_SLIP_PLANES = np.array(
    [[1, 1, 0], [1, 1, 0], [-1, 1, 0], [-1, 1, 0], [1, 0, 1], [1, 0, 1],
     [-1, 0, 1], [-1, 0, 1], [0, 1, 1], [0, 1, 1], [0, -1, 1], [0, -1, 1]], dtype=np.float64)
BCC_BURGERS = np.array(
    [[-1, 1, 1], [1, -1, 1], [1, 1, 1], [1, 1, -1], [1, 1, -1], [-1, 1, 1],
     [1, 1, 1], [1, 1, 1], [1, -1, 1], [1, 1, -1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)


@dataclass
class SlipSystem:
    """SlipSystem (DislocationType.jl:142-146): columns are systems."""

    slipPlane: np.ndarray  # 3 x nSys
    bVec: np.ndarray  # 3 x nSys

    @staticmethod
    def bcc() -> "SlipSystem":
        return SlipSystem(np.asfortranarray(BCC_SLIP_PLANES.T), np.asfortranarray(BCC_BURGERS.T))


def _norm(v):
    return math.sqrt(float(v[0]) * v[0] + float(v[1]) * v[1] + float(v[2]) * v[2])


def _cross(a, b):
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])


def externalAngle(n: int) -> float:
    """Misc.jl:143-151."""
    return math.pi - (n - 2) * math.pi / n


def rot3D(xyz, uvw, abc, θ):
    """Rotation of point xyz about the axis uvw through abc by θ (Misc.jl:179-207)."""
    uvw = np.asarray(uvw, dtype=np.float64)
    uvw = uvw / _norm(uvw)
    u1, u2, u3 = uvw
    a1, a2, a3 = abc
    x1, x2, x3 = xyz
    cosθ = math.cos(θ)
    onemcosθ = 1 - cosθ
    sinθ = math.sin(θ)
    d = x1 * u1 + x2 * u2 + x3 * u3
    return np.array([
        (a1 * (u2 * u2 + u3 * u3) - u1 * (a2 * u2 + a3 * u3 - d)) * onemcosθ + x1 * cosθ
        + sinθ * (-a3 * u2 + a2 * u3 - u3 * x2 + u2 * x3),
        (a2 * (u1 * u1 + u3 * u3) - u2 * (a1 * u1 + a3 * u3 - d)) * onemcosθ + x2 * cosθ
        + sinθ * (a3 * u1 - a1 * u3 + u3 * x1 - u1 * x3),
        (a3 * (u1 * u1 + u2 * u2) - u3 * (a1 * u1 + a2 * u2 - d)) * onemcosθ + x3 * cosθ
        + sinθ * (-a2 * u1 + a1 * u2 - u2 * x1 + u1 * x2),
    ])


@dataclass
class DislocationLoop:
    """A regular polygonal loop, built as DislocationLoop(loopPure, ...) does
    (DislocationConstructor.jl:196-298): b and n unit-normalised, first side
    along n x b, each next side rotated by the external angle about n (shear)
    or b (prism), re-centred on the mean node.

    ``dist`` is "zeros" (loops sit at range[:, 0]) or an explicit 3 x numLoops
    array of normalised displacements in [0, 1) (the role of ``rand(3, n)`` in
    loopDistribution, DislocationUtil.jl:96-104)."""

    loopType: str  # "shear" | "prism"
    numSides: int
    nodeSide: int
    numLoops: int
    segLen: Sequence[float]
    slipSystemIdx: int  # 1-based
    slipSystem: SlipSystem
    label: Sequence[int]
    buffer: float
    range: np.ndarray  # 3 x 2
    dist: object = "zeros"

    def __post_init__(self):
        nodeTotal = self.numSides * self.nodeSide
        segLen = np.asarray(self.segLen, dtype=np.float64).ravel()
        assert len(self.label) == nodeTotal
        assert len(segLen) == nodeTotal
        n = self.slipSystem.slipPlane[:, self.slipSystemIdx - 1].astype(np.float64)
        b = self.slipSystem.bVec[:, self.slipSystemIdx - 1].astype(np.float64)
        n = n / _norm(n)
        b = b / _norm(b)
        rotAxis = n if self.loopType == "shear" else b
        edge = _cross(n, b)
        edge = edge / _norm(edge)
        seg = [edge * L for L in segLen]
        θ = externalAngle(self.numSides)
        coord = np.zeros((3, nodeTotal))
        origin = (0.0, 0.0, 0.0)
        for i in range(1, self.numSides + 1):
            idx = (i - 1) * self.nodeSide
            modIdx = (i - 1) % len(segLen)
            rseg = rot3D(seg[modIdx], rotAxis, origin, θ * (i - 1))
            for j in range(1, self.nodeSide + 1):
                if i == j == 1:
                    coord[:, 0] = 0
                    continue
                if idx + j <= nodeTotal:
                    coord[:, idx + j - 1] += coord[:, idx + j - 2] + rseg
        mean = np.zeros(3)
        for k in range(nodeTotal):  # left-to-right sum, then divide
            mean = mean + coord[:, k]
        coord -= (mean / nodeTotal)[:, None]
        links = np.zeros((2, nodeTotal), dtype=np.int64)
        for j in range(1, nodeTotal):
            links[:, j - 1] = (j, j + 1)
        links[:, nodeTotal - 1] = (nodeTotal, 1)
        self.segLen = segLen
        self.links = links
        self.coord = coord
        self.bVec = np.repeat(b[:, None], nodeTotal, axis=1)
        self.slipPlane = np.repeat(n[:, None], nodeTotal, axis=1)
        self.label = np.asarray(self.label, dtype=np.int64)


@dataclass
class DislocationNetwork:
    """DislocationNetwork (DislocationType.jl:324-340).  numNode/numSeg are
    length-1 arrays as in the reference; buffers may be longer than the live
    part and kernels read only numSeg[0] segments."""

    numNode: np.ndarray
    numSeg: np.ndarray
    maxConnect: int
    label: np.ndarray
    links: np.ndarray  # 2 x buf, int64 1-based
    connectivity: np.ndarray  # (1+2 maxConnect) x buf
    linksConnect: np.ndarray
    slipPlane: np.ndarray
    extSeg: np.ndarray  # bool
    segIdx: np.ndarray  # buf x 3, int64 1-based, F-order
    bVec: np.ndarray  # 3 x buf
    coord: np.ndarray  # 3 x buf
    nodeVel: np.ndarray
    nodeForce: np.ndarray
    segForce: np.ndarray  # 3 x 2 x buf, F-order

    @staticmethod
    def from_loops(sources, maxConnect: int = 4, memBuffer: Optional[int] = None):
        """DislocationNetwork(sources; memBuffer) (DislocationConstructor.jl:477-551)
        with makeNetwork! (:645-688)."""
        if isinstance(sources, DislocationLoop):
            sources = [sources]
        nodeTotal = sum(s.numLoops * len(s.label) for s in sources)
        nodeBuffer = (int(round(nodeTotal * math.log2(nodeTotal))) if memBuffer is None
                      else nodeTotal * int(memBuffer))
        F = dict(order="F")
        links = np.zeros((2, nodeBuffer), dtype=np.int64, **F)
        slipPlane = np.zeros((3, nodeBuffer), **F)
        bVec = np.zeros((3, nodeBuffer), **F)
        coord = np.zeros((3, nodeBuffer), **F)
        label = np.zeros(nodeBuffer, dtype=np.int64)
        total = 0
        for s in sources:
            nl = len(s.label)
            if isinstance(s.dist, str) and s.dist == "zeros":
                disp = np.zeros((3, s.numLoops))
            else:
                disp = np.asarray(s.dist, dtype=np.float64).reshape(3, s.numLoops)
            lims = np.asarray(s.range, dtype=np.float64) + s.buffer * float(np.mean(s.segLen))
            for j in range(s.numLoops):
                a, b = total, total + nl
                links[:, a:b] = s.links + total
                slipPlane[:, a:b] = s.slipPlane
                bVec[:, a:b] = s.bVec
                label[a:b] = s.label
                shift = lims[:, 0] + (lims[:, 1] - lims[:, 0]) * disp[:, j]
                coord[:, a:b] = s.coord + shift[:, None]
                total += nl
        return DislocationNetwork._assemble(links, slipPlane, bVec, coord, label, nodeTotal,
                                            nodeTotal, maxConnect)

    @staticmethod
    def _assemble(links, slipPlane, bVec, coord, label, numNode, numSeg, maxConnect):
        buf = links.shape[1]
        segIdx, extSeg = getSegmentIdx(links, label)
        connectivity, linksConnect = makeConnect(links, maxConnect, coord.shape[1])
        F = dict(order="F")
        return DislocationNetwork(
            numNode=np.array([numNode], dtype=np.int64), numSeg=np.array([numSeg], dtype=np.int64),
            maxConnect=maxConnect, label=label, links=links, connectivity=connectivity,
            linksConnect=linksConnect, slipPlane=slipPlane, extSeg=extSeg, segIdx=segIdx,
            bVec=bVec, coord=coord, nodeVel=np.zeros((3, coord.shape[1]), **F),
            nodeForce=np.zeros((3, coord.shape[1]), **F),
            segForce=np.zeros((3, 2, buf), **F))

    @staticmethod
    def from_arrays(links, bVec, coord, label, maxConnect: int = 4, slipPlane=None):
        """Keyword DislocationNetwork constructor followed by makeConnect! and
        getSegmentIdx! (DislocationConstructor.jl:412-455), as the
        field-stress tests build their one-segment networks."""
        links = np.asfortranarray(links, dtype=np.int64)
        bVec = np.asfortranarray(bVec, dtype=np.float64)
        coord = np.asfortranarray(coord, dtype=np.float64)
        label = np.asarray(label, dtype=np.int64)
        if slipPlane is None:
            slipPlane = np.zeros_like(bVec, order="F")
        return DislocationNetwork._assemble(links, slipPlane, bVec, coord, label, len(label),
                                            links.shape[1], maxConnect)


def makeConnect(links, maxConnect, numNodeCols=None):
    """makeConnect (DislocationUtil.jl:193-237): per node, the count of
    connections then (link, 1|2) pairs; per link, its slot in each node."""
    lenLinks = links.shape[1]
    ncols = lenLinks if numNodeCols is None else max(lenLinks, numNodeCols)
    connectivity = np.zeros((1 + 2 * maxConnect, ncols), dtype=np.int64, order="F")
    linksConnect = np.zeros((2, lenLinks), dtype=np.int64, order="F")
    zero = np.nonzero(links[0, :] == 0)[0]
    n = lenLinks if len(zero) == 0 else int(zero[0])
    for i in range(n):
        n1, n2 = links[0, i] - 1, links[1, i] - 1
        connectivity[0, n1] += 1
        connectivity[0, n2] += 1
        t1 = 2 * connectivity[0, n1]
        t2 = 2 * connectivity[0, n2]
        connectivity[t1 - 1, n1], connectivity[t1, n1] = i + 1, 1
        connectivity[t2 - 1, n2], connectivity[t2, n2] = i + 1, 2
        linksConnect[0, i] = connectivity[0, n1]
        linksConnect[1, i] = connectivity[0, n2]
    return connectivity, linksConnect


def getSegmentIdx(links, label):
    """getSegmentIdx (DislocationUtil.jl:253-295): segIdx[i, :] = (i, n1, n2);
    a segment touching an external node is external."""
    lenLinks = links.shape[1]
    segIdx = np.zeros((lenLinks, 3), dtype=np.int64, order="F")
    extSeg = np.zeros(lenLinks, dtype=bool)
    zero = np.nonzero(links[0, :] == 0)[0]
    n = lenLinks if len(zero) == 0 else int(zero[0])
    i = np.arange(n)
    segIdx[:n, 0] = i + 1
    segIdx[:n, 1] = links[0, :n]
    segIdx[:n, 2] = links[1, :n]
    lab = np.asarray(label)
    extSeg[:n] = (lab[links[0, :n] - 1] == extDln) | (lab[links[1, :n] - 1] == extDln)
    return segIdx, extSeg


@dataclass
class RegularCuboidMesh:
    """The parts of RegularCuboidMesh (FEMType.jl:212-272) that calc_σHat reads,
    built as buildMesh does for linear elements (FEMConstructor.jl:126-496):
    isotropic stiffness C (Voigt order xx, yy, zz, xy, xz, yz), node
    coordinates, 8-node element connectivity, element counts and sizes."""

    dx: float
    dy: float
    dz: float
    mx: int
    my: int
    mz: int
    μ: float
    ν: float
    C: np.ndarray = field(init=False)
    coord: np.ndarray = field(init=False)
    connectivity: np.ndarray = field(init=False)

    def __post_init__(self):
        mx, my, mz = int(self.mx), int(self.my), int(self.mz)
        self.w, self.h, self.d = self.dx / mx, self.dy / my, self.dz / mz
        μ, ν = float(self.μ), float(self.ν)
        E = 2 * μ * (1 + ν)
        νopνInv = ν * (1 / (1 + ν))
        Lame = E * νopνInv / (1 - 2 * ν)
        CDiag = Lame + 2 * μ
        C = np.zeros((6, 6), order="F")
        C[:3, :3] = Lame
        C[0, 0] = C[1, 1] = C[2, 2] = CDiag
        C[3, 3] = C[4, 4] = C[5, 5] = μ
        self.C = C
        mx1, my1, mz1 = mx + 1, my + 1, mz + 1
        self.numNode = mx1 * my1 * mz1
        self.numElem = mx * my * mz
        coord = np.zeros((3, self.numNode), order="F")
        for k in range(1, mz1 + 1):
            for j in range(1, my1 + 1):
                for i in range(1, mx1 + 1):
                    g = i + (j - 1) * mx1 + (k - 1) * mx1 * my1
                    coord[:, g - 1] = ((i - 1) * self.w, (j - 1) * self.h, (k - 1) * self.d)
        conn = np.zeros((8, self.numElem), dtype=np.int64, order="F")
        for k in range(1, mz + 1):
            for j in range(1, my + 1):
                for i in range(1, mx + 1):
                    e = i + (j - 1) * mx + (k - 1) * mx * my - 1
                    c1 = i + (j - 1) * mx1 + (k - 1) * mx1 * my1
                    c4 = i + j * mx1 + (k - 1) * mx1 * my1
                    c5 = i + (j - 1) * mx1 + k * mx1 * my1
                    c8 = i + j * mx1 + k * mx1 * my1
                    conn[:, e] = (c1, c1 + 1, c4 + 1, c4, c5, c5 + 1, c8 + 1, c8)
        self.coord = coord
        self.connectivity = conn
        # Bounding polytope (FEMConstructor.jl:164-236): the 8 corners, the 6
        # faces by corner number, their midpoints and outward normals.
        dx, dy, dz = float(self.dx), float(self.dy), float(self.dz)
        self.vertices = np.asfortranarray(np.array(
            [[0, 0, 0], [dx, 0, 0], [dx, dy, 0], [0, dy, 0],
             [0, 0, dz], [dx, 0, dz], [dx, dy, dz], [0, dy, dz]], dtype=np.float64).T)
        self.faces = np.asfortranarray(np.array(
            [[1, 2, 6, 5], [2, 3, 7, 6], [3, 4, 8, 7], [4, 1, 5, 8], [4, 3, 2, 1], [5, 6, 7, 8]],
            dtype=np.int64).T)
        self.faceMidPt = np.asfortranarray(
            self.vertices[:, self.faces - 1].sum(axis=1) / 4)
        self.faceNorm = np.asfortranarray(np.array(
            [[0, -1, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, 0, -1], [0, 0, 1]],
            dtype=np.float64).T)
        self.surfNode = _surface_nodes(mx, my, mz)


def _surface_nodes(mx, my, mz):
    """surfNode of buildMesh (FEMConstructor.jl:274-305,499-545): node
    numbers of the 8 corners, the edge nodes along each of the 12 edges and
    the interior nodes of each of the 6 faces."""
    mx1, my1, mz1 = mx + 1, my + 1, mz + 1
    s = {"x0y0z0": [1], "x1y0z0": [mx1], "x1y1z0": [mx1 * my1], "x0y1z0": [mx1 * my + 1],
         "x0y0z1": [1 + mx1 * my1 * mz], "x1y0z1": [mx1 + mx1 * my1 * mz],
         "x1y1z1": [mx1 * my1 * mz1], "x0y1z1": [mx1 * my + 1 + mx1 * my1 * mz]}
    top = mx1 * my1 * mz
    s["x_y0z0"] = [i + 1 for i in range(1, mx)]
    s["x_y1z0"] = [i + 1 + mx1 * my for i in range(1, mx)]
    s["x_y0z1"] = [i + 1 + top for i in range(1, mx)]
    s["x_y1z1"] = [i + 1 + top + mx1 * my for i in range(1, mx)]
    s["y_x0z0"] = [i * mx1 + 1 for i in range(1, my)]
    s["y_x1z0"] = [i * mx1 + mx1 for i in range(1, my)]
    s["y_x0z1"] = [i * mx1 + 1 + top for i in range(1, my)]
    s["y_x1z1"] = [i * mx1 + mx1 + top for i in range(1, my)]
    s["z_x0y0"] = [i * mx1 * my1 + 1 for i in range(1, mz)]
    s["z_x1y0"] = [i * mx1 * my1 + mx1 for i in range(1, mz)]
    s["z_x0y1"] = [i * mx1 * my1 + 1 + mx1 * my for i in range(1, mz)]
    s["z_x1y1"] = [i * mx1 * my1 + mx1 + mx1 * my for i in range(1, mz)]
    s["xy_z0"] = [j * mx1 + 1 + i for j in range(1, my) for i in range(1, mx)]
    s["xy_z1"] = [j * mx1 + 1 + i + top for j in range(1, my) for i in range(1, mx)]
    s["xz_y0"] = [j * mx1 * my1 + 1 + i for j in range(1, mz) for i in range(1, mx)]
    s["xz_y1"] = [j * mx1 * my1 + 1 + i + mx1 * my for j in range(1, mz) for i in range(1, mx)]
    s["yz_x0"] = [j * mx1 * my1 + 1 + i * mx1 for j in range(1, mz) for i in range(1, my)]
    s["yz_x1"] = [j * mx1 * my1 + mx1 + i * mx1 for j in range(1, mz) for i in range(1, my)]
    return {k: np.asarray(v, dtype=np.int64) for k, v in s.items()}


@dataclass
class Boundaries:
    """The node and dof lists of Boundaries (FEMType.jl:330-370) that
    calc_uTilde reads: uGammaDln = [uGamma.node; mGamma.node] and
    uDofsDln = [3 uGammaDln - 2; 3 uGammaDln - 1; 3 uGammaDln]
    (FEMConstructor.jl:968-971: all x dofs, then all y, then all z)."""

    uGammaDln: np.ndarray
    uDofsDln: np.ndarray = field(init=False)

    def __post_init__(self):
        u = np.asarray(self.uGammaDln, dtype=np.int64)
        self.uGammaDln = u
        self.uDofsDln = np.concatenate([3 * u - 2, 3 * u - 1, 3 * u])

    @staticmethod
    def cantilever(mesh: "RegularCuboidMesh") -> "Boundaries":
        """Boundaries for CantileverLoad (FEMConstructor.jl:754-800): the x = 0
        face with its corners and edges (uGamma), and the loaded corners and
        edge at x = dx, z = dz (mGamma)."""
        sn = mesh.surfNode
        uIndex = ["x0y0z0", "x0y1z0", "x0y0z1", "x0y1z1", "y_x0z0", "y_x0z1", "z_x0y0",
                  "z_x0y1", "yz_x0"]
        mIndex = ["x1y0z1", "x1y1z1", "y_x1z1"]
        return Boundaries(np.concatenate([sn[k] for k in uIndex + mIndex]))


@dataclass
class ForceDisplacement:
    """ForceDisplacement (FEMType.jl:294-301): ``uHat`` (read by the force
    path) and ``uTilde`` (written by calc_uTilde!).  The reference stores
    them as SparseVectors; here they are dense (3 numNode)."""

    uHat: np.ndarray
    uTilde: Optional[np.ndarray] = None
