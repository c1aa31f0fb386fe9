"""Error measures used by the parity tests."""
import numpy as np


def per_node_rel(test, ref, tau=1e-6):
    """Per-node relative error |dF| / max(|F|, tau * RMS|F|) over 3-vectors.

    ``test`` and ``ref`` are arrays whose first axis is the 3 Cartesian
    components (3 x n or 3 x 2 x n).  The floor tau * RMS|F| keeps nodes whose
    force nearly cancels from dominating the measure.  Returns (max, rms)."""
    a = np.asarray(test, dtype=np.float64).reshape(3, -1, order="F")
    b = np.asarray(ref, dtype=np.float64).reshape(3, -1, order="F")
    nb = np.linalg.norm(b, axis=0)
    rms = np.sqrt(np.mean(nb ** 2)) if nb.size else 0.0
    den = np.maximum(nb, tau * rms)
    den[den == 0] = 1.0
    r = np.linalg.norm(a - b, axis=0) / den
    return float(r.max(initial=0.0)), float(np.sqrt(np.mean(r ** 2))) if r.size else 0.0


def normwise(test, ref):
    """Julia's isapprox measure: ||x - y|| / max(||x||, ||y||)."""
    a = np.asarray(test, dtype=np.float64).ravel()
    b = np.asarray(ref, dtype=np.float64).ravel()
    d = max(np.linalg.norm(a), np.linalg.norm(b))
    return float(np.linalg.norm(a - b) / d) if d > 0 else 0.0


def isapprox(test, ref, rtol=np.sqrt(np.finfo(np.float64).eps)):
    return normwise(test, ref) <= rtol


def node_forces(net, sf):
    """Segment-end forces 3 x 2 x numSeg gathered to the nodes they act on
    (3 x nodes used), as dlnMobility!'s node-force gather adds them."""
    n = sf.shape[2]
    si = net.segIdx[:n]
    F = np.zeros((net.coord.shape[1], 3))
    np.add.at(F, si[:, 1] - 1, np.asarray(sf)[:, 0, :].T)
    np.add.at(F, si[:, 2] - 1, np.asarray(sf)[:, 1, :].T)
    used = np.unique(si[:, 1:3]) - 1
    return F[used].T
