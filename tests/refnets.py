"""The networks, meshes and parameters of the reference's force-path tests,
rebuilt from the committed golden data.

Each builder names the reference test it follows; the expected outputs live in
tests/golden/reference_vectors.json.
"""
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "ddd.jl_amd"))
sys.path.insert(0, str(ROOT / "oracle"))

from dddhip.types import (DislocationLoop, DislocationNetwork, DislocationParameters,  # noqa: E402
                          MaterialParameters, RegularCuboidMesh, ForceDisplacement, SlipSystem,
                          intMobDln, intFixDln)

GOLD = json.loads((ROOT / "tests/golden/reference_vectors.json").read_text())


def g(name):
    return np.asarray(GOLD[name]["data"], dtype=np.float64)


def params():
    """loadParameters of inputs/simParams/sample{Dislocation,Material}Parameters.json
    (Input.jl:52-64,244-269): coreRad 90.78, Ec 1, mu 1, nu 0.28, sigmaPN 0,
    drag edge 1, screw 2, climb 1e10, line 0."""
    d = GOLD["file:inputs/simParams/sampleDislocationParameters.json"]["data"]
    m = GOLD["file:inputs/simParams/sampleMaterialParameters.json"]["data"]
    dln = DislocationParameters(coreRad=float(d["coreRad"]), coreEnergy=float(d["coreEnergy"]),
                                dragCoeffs={k: float(v) for k, v in d["dragCoeffs"].items()})
    mat = MaterialParameters(μ=float(m["μ"]), ν=float(m["ν"]), μMag=float(m["μMag"]),
                             σPN=float(m["σPN"]))
    return dln, mat


def integration():
    """loadIntegrationParameters / loadIntegrationTime of the sample files
    (Input.jl:196-213,352-360).  maxiter is read from the "exponent" key
    (Input.jl:209), so the sample's maxiter is 20, not 10."""
    from dddhip.types import IntegrationParameters, IntegrationTime
    d = GOLD["file:inputs/simParams/sampleIntegrationParameters.json"]["data"]
    t = GOLD["file:test/testData/sampleIntegrationTime.json"]["data"]
    ip = IntegrationParameters(tmin=float(d["tmin"]), tmax=float(d["tmax"]),
                               dtmin=float(d["dtmin"]), dtmax=float(d["dtmax"]),
                               abstol=float(d["abstol"]), reltol=float(d["reltol"]),
                               maxchange=float(d["maxchange"]), exponent=float(d["exponent"]),
                               maxiter=int(d["exponent"]))
    return ip, IntegrationTime(dt=float(t["dt"]), time=float(t["time"]), step=int(t["step"]))


def utilde_setup():
    """The calc_uTilde test of DislocationFEMTest.jl:270-320: a 1013 x 1987
    x 2999 cantilever meshed 3 x 5 x 7, a prismatic and a shear octagon on
    the slip system (-1 1 0)[1 1 1], both centred in the volume, nu = 0.28."""
    from dddhip.types import Boundaries, ForceDisplacement
    dx, dy, dz = 1013.0, 1987.0, 2999.0
    mat = MaterialParameters(μ=1.0, ν=0.28)
    m = RegularCuboidMesh(dx=dx, dy=dy, dz=dz, mx=3, my=5, mz=7, μ=mat.μ, ν=mat.ν)
    ss = SlipSystem(np.array([[-1.0], [1.0], [0.0]]), np.array([[1.0], [1.0], [1.0]]))
    segLen = (dx + dy + dz) / 30
    rng = np.array([[dx / 2, dx / 2], [dy / 2, dy / 2], [dz / 2, dz / 2]])

    def loop(kind):
        return DislocationLoop(loopType=kind, numSides=8, nodeSide=1, numLoops=1,
                               segLen=segLen * np.ones(8), slipSystemIdx=1, slipSystem=ss,
                               label=[intMobDln] * 8, buffer=0.0, range=rng, dist="zeros")
    net = DislocationNetwork.from_loops([loop("prism"), loop("shear")])
    bnd = Boundaries.cantilever(m)
    fd = ForceDisplacement(uHat=np.zeros(3 * m.numNode), uTilde=np.zeros(3 * m.numNode))
    return mat, m, bnd, fd, net


def collision_params():
    """DislocationParameters of test/testData/sampleDislocationParameters.json
    (TopologyTest.jl:6-20): collision distance 157.235."""
    d = GOLD["file:test/testData/sampleDislocationParameters.json"]["data"]
    return DislocationParameters(coreRad=float(d["coreRad"]), coreEnergy=float(d["coreEnergy"]),
                                 collisionDist=float(d["collisionDist"]))


def collision_network(k):
    """Network k of TopologyTest.jl's collision checks: DislocationNetwork
    with the given links, coordinates, labels and node velocities, then
    makeConnect! and getSegmentIdx!."""
    d = GOLD["collision_cases"]["data"]["networks"][k]
    links = np.asarray(d["links"], dtype=np.int64)
    net = DislocationNetwork.from_arrays(links, np.asarray(d["bVec"]),
                                         np.asarray(d["coord"]),
                                         np.asarray(d["label"]).ravel().astype(np.int64))
    net.nodeVel[:, :links.max()] = np.asarray(d["nodeVel"])
    return net


def mesh():
    """buildMesh(matParams, femParams) for sampleFEMParameters.json: 2000^3, 5x5x5."""
    f = GOLD["file:test/testData/sampleFEMParameters.json"]["data"]
    _, mat = params()
    return RegularCuboidMesh(dx=f["dx"], dy=f["dy"], dz=f["dz"], mx=f["mx"], my=f["my"],
                             mz=f["mz"], μ=mat.μ, ν=mat.ν)


def sparse_uhat(m, index_key, value_key, scale=1000.0):
    u = np.zeros(3 * m.numNode)
    idx = np.asarray(GOLD[index_key]["data"], dtype=np.int64)
    u[idx - 1] = g(value_key) * scale
    return ForceDisplacement(uHat=u)


def _loop(kind, sides, sys_idx, labels, numLoops=1, seglen=10.0):
    return DislocationLoop(loopType=kind, numSides=sides, nodeSide=1, numLoops=numLoops,
                           segLen=seglen * np.ones(sides), slipSystemIdx=sys_idx,
                           slipSystem=SlipSystem.bcc(), label=labels, buffer=0.0,
                           range=np.array([[-100.0, 100.0]] * 3), dist="zeros")


def pentagons(shift=0.0):
    """Two prismatic pentagons on slip system 4, segLen 10, Zeros(), the second
    shifted by +shift (SegmentForcesTest.jl:221-234,281-294,326-340)."""
    lp = _loop("prism", 5, 4, [intMobDln, intFixDln, intMobDln, intFixDln, intMobDln], 2)
    net = DislocationNetwork.from_loops(lp, memBuffer=1)
    net.coord[:, 5:] += shift
    return net


def hexagons():
    """Prismatic + shear hexagons on slip system 1 (SegmentForcesTest.jl:382-409)."""
    lab = [intMobDln, intFixDln] * 3
    return DislocationNetwork.from_loops([_loop("prism", 6, 1, lab), _loop("shear", 6, 1, lab)],
                                         memBuffer=1)


def mobility_network():
    """Shear hexagon + prismatic pentagon on slip system 4 (MobilityTest.jl:21-49)."""
    pent = _loop("prism", 5, 4, [intMobDln, intFixDln, intMobDln, intFixDln, intMobDln])
    hexa = _loop("shear", 6, 4, [intMobDln, intFixDln, intMobDln, intFixDln, intMobDln,
                                 intMobDln])
    return DislocationNetwork.from_loops([hexa, pent], memBuffer=1)


def pk_network():
    """The 32-node network of the PK test: two prismatic and two shear 8-node
    squares on slip system 1 with explicit coordinates
    (SegmentForcesTest.jl:83-130)."""
    m = mesh()
    segLen = m.dx / 8
    lab = [intMobDln] * 8
    rng = np.array([[segLen, m.dx - segLen], [segLen, m.dy - segLen], [segLen, m.dz - segLen]])

    def sq(kind):
        return DislocationLoop(loopType=kind, numSides=4, nodeSide=2, numLoops=2,
                               segLen=segLen * np.ones(8), slipSystemIdx=1,
                               slipSystem=SlipSystem.bcc(), label=lab, buffer=0.0, range=rng,
                               dist=np.full((3, 2), 0.5))
    net = DislocationNetwork.from_loops([sq("prism"), sq("shear")])
    n = int(net.numSeg[0])
    net.coord[:, :n] = g("pk_coord")
    return net


def sigma_segment(b):
    """One segment from (0,0,-300) to (0,0,0), core radius 5, and the 3x3 grid of
    points on z = 0 (DislocationFEMTest.jl:6-57)."""
    xr = np.linspace(-300, 300, 3)
    X = np.ones(3)[:, None] * xr[None, :]
    Y = xr[:, None] * np.ones(3)[None, :]
    points = np.asfortranarray(np.vstack([X.ravel(order="F"), Y.ravel(order="F"),
                                          np.zeros(9)]))
    links = np.array([[1], [2]])
    bVec = np.asarray(b, dtype=np.float64).reshape(3, 1)
    coord = np.vstack([np.zeros(3), np.zeros(3), xr])
    net = DislocationNetwork.from_arrays(links, bVec, coord, [0, 0, 0])
    mat = MaterialParameters(μ=1.0, ν=0.28)
    dln = DislocationParameters(coreRad=5.0)
    return points, dln, mat, net
