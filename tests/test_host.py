"""CPU tests of the host side: the C-ABI library loads and exports every entry
point include/ddd_hip.h declares, the API fails loudly without a device, and
the network construction mirrors the reference's conventions."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from dddhip import _lib as L
from dddhip import forces as G
from dddhip.generate import splitmix64_uniform, random_loop_network, face_grid_points
from dddhip.types import (DislocationNetwork, makeConnect, getSegmentIdx, extDln,
                          DislocationParameters, MaterialParameters)

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "ddd_hip.h").read_text()
    return sorted(set(re.findall(r"\b(ddd_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(L.LIB_PATH))
    names = _declared()
    assert "ddd_segseg_force" in names and "ddd_sigma_tilde" in names
    for n in names:
        assert hasattr(lib, n), n
    assert L.lib().ddd_version() == 1


def _device_count():
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return n.value if hip.hipGetDeviceCount(ctypes.byref(n)) == 0 else 0
    except OSError:
        return 0


@pytest.mark.skipif(_device_count() > 0, reason="a GPU is present")
def test_no_device_fails_loudly():
    with pytest.raises(L.DDDError) as e:
        L.Context(0)
    assert e.value.code == L.DDD_ERR_NO_DEVICE
    net = random_loop_network(2, seed=1)
    dln, mat = DislocationParameters(coreRad=90.78, coreEnergy=1.0), MaterialParameters(ν=0.28)
    with pytest.raises(L.DDDError):
        G.calcSegSegForce(dln, mat, net)


def test_null_context_is_an_argument_error():
    lib = L.lib()
    assert lib.ddd_segseg_force(None, None, None, None, 0, None, 0, 0) == L.DDD_ERR_ARG
    assert lib.ddd_net_seg_force(None, 0) == L.DDD_ERR_ARG
    assert lib.ddd_ctx_destroy(None) == L.DDD_OK


def test_splitmix64_reference_values():
    # SplitMix64 (Steele, Lea & Flood 2014) from state 0: first outputs
    # 0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4.
    u = splitmix64_uniform(0, 2)
    assert u[0] == (0xE220A8397B1DCDAF >> 11) * 2.0 ** -53
    assert u[1] == (0x6E789E6AA1B965F4 >> 11) * 2.0 ** -53


def test_generator_deterministic_and_shaped():
    a = random_loop_network(5, seed=0xA0)
    b = random_loop_network(5, seed=0xA0)
    assert np.array_equal(a.coord, b.coord) and np.array_equal(a.bVec, b.bVec)
    n = int(a.numSeg[0])
    assert n == 160 and int(a.numNode[0]) == 160
    # closed 32-gons: side length 2 R sin(pi/32) with R in [100, 300]
    L = np.linalg.norm(a.coord[:, a.segIdx[:n, 2] - 1] - a.coord[:, a.segIdx[:n, 1] - 1], axis=0)
    assert np.all((L > 2 * 100 * np.sin(np.pi / 32) - 1e-9) & (L < 2 * 300 * np.sin(np.pi / 32)))
    # unit Burgers vectors of the 12 BCC systems, perpendicular to the loop plane normals
    assert np.allclose(np.linalg.norm(a.bVec[:, :n], axis=0), 1.0)
    assert np.allclose(np.einsum("ij,ij->j", a.bVec[:, :n], a.slipPlane[:, :n]), 0.0)


def test_connectivity_conventions():
    # one triangle loop: links (1,2), (2,3), (3,1)
    links = np.array([[1, 2, 3], [2, 3, 1]], dtype=np.int64)
    conn, lc = makeConnect(links, 4)
    assert conn[:, 0].tolist()[:5] == [2, 1, 1, 3, 2]  # node 1: link 1 as node 1, link 3 as node 2
    assert conn[:, 1].tolist()[:5] == [2, 1, 2, 2, 1]
    assert lc[:, 0].tolist() == [1, 1]
    seg, ext = getSegmentIdx(links, np.array([1, 1, extDln]))
    assert seg.tolist() == [[1, 1, 2], [2, 2, 3], [3, 3, 1]]
    assert ext.tolist() == [False, True, True]


def test_face_grid_points():
    p = face_grid_points(128)
    assert p.shape == (3, 98304)
    assert p.min() == 0.0 and p.max() == 2000.0
    on_face = (p == 0.0) | (p == 2000.0)
    assert np.all(on_face.sum(axis=0) == 1)


def test_network_buffers_mirror_reference():
    net = random_loop_network(3, seed=7, memBuffer=None)
    n = int(net.numSeg[0])
    assert net.segForce.shape[2] == round(n * np.log2(n))  # DislocationConstructor.jl:493-494
    assert net.segIdx.flags.f_contiguous and net.segForce.flags.f_contiguous
    assert isinstance(net, DislocationNetwork)
