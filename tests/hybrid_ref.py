"""numpy restatement of the hybrid far field (test infrastructure): cell
binning, cell moments, the direct (non-FFT) correlation with an operator W,
and the local-expansion quadrature onto segments."""
import numpy as np

KK = [0, 1, 2, 0, 0, 1]
LL = [0, 1, 2, 1, 2, 2]


def segments(net):
    n = int(net.numSeg[0])
    si = net.segIdx[:n]
    n1 = net.coord[:, si[:, 1] - 1].T
    n2 = net.coord[:, si[:, 2] - 1].T
    b = net.bVec[:, si[:, 0] - 1].T
    return n1, n2, b, net.extSeg[:n].astype(bool)


def bin_cells(net, G, lo, h):
    n1, n2, _, ext = segments(net)
    mid = 0.5 * (n1 + n2)
    c3 = np.clip(np.floor((mid - np.asarray(lo)) / h).astype(np.int64), 0, G - 1)
    cell = c3[:, 0] + G * (c3[:, 1] + G * c3[:, 2])
    cell[ext] = -1
    return cell


def _spread(v):
    r = np.zeros_like(v)
    for b in range(10):
        r |= ((v >> b) & 1) << (3 * b)
    return r


def sorted_rank(net, G, lo, h):
    """Position of each binned segment in the near-field order: cells by the
    Morton code of (cx, cy, cz), then segment index; -1 for unbinned."""
    cell = bin_cells(net, G, lo, h)
    ok = cell >= 0
    c = np.where(ok, cell, 0).astype(np.int64)
    code = _spread(c % G) | (_spread((c // G) % G) << 1) | (_spread(c // (G * G)) << 2)
    idx = np.arange(len(cell))
    o = np.lexsort((idx[ok], code[ok]))
    rank = np.full(len(cell), -1, dtype=np.int64)
    rank[idx[ok][o]] = np.arange(ok.sum())
    return rank


def moments(net, G, lo, h):
    n1, n2, b, ext = segments(net)
    cell = bin_cells(net, G, lo, h)
    nc = G ** 3
    M = np.zeros((90, nc))
    for i in np.nonzero(cell >= 0)[0]:
        c = cell[i]
        cc = np.asarray(lo) + (np.array([c % G, (c // G) % G, c // (G * G)]) + 0.5) * h
        l = (n2[i] - n1[i]) / h
        r = (0.5 * (n1[i] + n2[i]) - cc) / h
        f = [b[i][a] * l[j] for a in range(3) for j in range(3)]
        f += [b[i][a] * l[j] * r[k] for a in range(3) for j in range(3) for k in range(3)]
        f += [b[i][a] * l[j] * (r[KK[m]] * r[LL[m]] + l[KK[m]] * l[LL[m]] / 12)
              for a in range(3) for j in range(3) for m in range(6)]
        M[:, c] += f
    return M


def local_expansions(M, W, G, rn):
    """L[:, t] = sum over far o of W(o) M[:, t + o], W indexed by
    (ox + G - 1) + S (oy + G - 1) + S^2 (oz + G - 1), S = 2G - 1."""
    S = 2 * G - 1
    nc = G ** 3
    L = np.zeros((60, nc))
    W = W.reshape(S ** 3, 60, 90)
    for t in range(nc):
        tx, ty, tz = t % G, (t // G) % G, t // (G * G)
        for sc in range(nc):
            sx, sy, sz = sc % G, (sc // G) % G, sc // (G * G)
            o = (sx - tx, sy - ty, sz - tz)
            if max(abs(o[0]), abs(o[1]), abs(o[2])) <= rn:
                continue
            oi = (o[0] + G - 1) + S * ((o[1] + G - 1) + S * (o[2] + G - 1))
            L[:, t] += W[oi] @ M[:, sc]
    return L


def basis(p):
    x, y, z = p
    return np.array([1.0, x, y, z, x * x, y * y, z * z, x * y, x * z, y * z])


def far_forces(net, L, G, lo, h):
    """Two-point quadrature of (sigma b) x t over each segment, sigma from its
    cell's expansion: 3 x 2 x numSeg."""
    n1, n2, b, ext = segments(net)
    cell = bin_cells(net, G, lo, h)
    out = np.zeros((3, 2, len(cell)), order="F")
    gq = [0.21132486540518711775, 0.78867513459481288225]
    for i in np.nonzero(cell >= 0)[0]:
        c = cell[i]
        cc = np.asarray(lo) + (np.array([c % G, (c // G) % G, c // (G * G)]) + 0.5) * h
        d = n2[i] - n1[i]
        Ln = np.linalg.norm(d)
        t = d / Ln
        f = []
        for u in gq:
            P = basis((n1[i] + u * d - cc) / h)
            s = L[:, c].reshape(6, 10) @ P
            S = np.array([[s[0], s[3], s[4]], [s[3], s[1], s[5]], [s[4], s[5], s[2]]])
            f.append(np.cross(S @ b[i], t) * (0.5 * Ln))
        out[:, 0, i] = f[0] * gq[1] + f[1] * gq[0]
        out[:, 1, i] = f[0] * gq[0] + f[1] * gq[1]
    return out
