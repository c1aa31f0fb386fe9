"""The cubic symmetry of the hybrid far field, restated for the tests: the 48
signed permutations R of the axes, (R v)[perm[k]] = sgn[k] v[k], and the
signed index permutations T_M(R) of the 90 cell moments and T_L(R) of the 60
local-expansion entries (layouts of tests/hybrid_ref.py: moments 3a + j,
9 + 9a + 3j + k, 36 + 18a + 6j + p; locals 10 c + l with the pairs xx, yy,
zz, xy, xz, yz and the basis 1, x, y, z, xx, yy, zz, xy, xz, yz)."""
import itertools

import numpy as np

PAIRS = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]


def pair_index(i, j):
    return PAIRS.index((min(i, j), max(i, j)))


def maps():
    out = []
    for perm in itertools.permutations(range(3)):
        for sg in itertools.product((1, -1), repeat=3):
            out.append((perm, sg))
    return out


def det(R):
    """+1 for rotations, -1 for reflections.  A reflection also reverses the
    right-handed Burgers circuit, so the stress of the reflected network is
    det(R) R sigma R^T and W(R o) = det(R) T_L(R) W(o) T_M(R)^T."""
    return int(round(np.linalg.det(matrix(R))))


def matrix(R):
    perm, sg = R
    A = np.zeros((3, 3))
    for k in range(3):
        A[perm[k], k] = sg[k]
    return A


def T_moments(R):
    perm, sg = R
    T = np.zeros((90, 90))
    for a in range(3):
        for j in range(3):
            T[3 * perm[a] + perm[j], 3 * a + j] = sg[a] * sg[j]
            for m in range(3):
                T[9 + 9 * perm[a] + 3 * perm[j] + perm[m], 9 + 9 * a + 3 * j + m] = \
                    sg[a] * sg[j] * sg[m]
            for p, (u, v) in enumerate(PAIRS):
                T[36 + 18 * perm[a] + 6 * perm[j] + pair_index(perm[u], perm[v]),
                  36 + 18 * a + 6 * j + p] = sg[a] * sg[j] * sg[u] * sg[v]
    return T


def T_locals(R):
    perm, sg = R
    T = np.zeros((60, 60))
    for c, (i, j) in enumerate(PAIRS):
        c2, sc = pair_index(perm[i], perm[j]), sg[i] * sg[j]
        T[10 * c2, 10 * c] = sc
        for k in range(3):
            T[10 * c2 + 1 + perm[k], 10 * c + 1 + k] = sc * sg[k]
        for p, (u, v) in enumerate(PAIRS):
            T[10 * c2 + 4 + pair_index(perm[u], perm[v]), 10 * c + 4 + p] = sc * sg[u] * sg[v]
    return T


def rotate_network(net, R, centre):
    """A copy of net with every node moved to centre + R (x - centre) and
    every Burgers vector to R b."""
    import copy
    A = matrix(R)
    out = copy.deepcopy(net)
    c = np.asarray(centre, dtype=np.float64).reshape(3, 1)
    out.coord[:] = c + A @ (net.coord - c)
    out.bVec[:] = A @ net.bVec
    return out
