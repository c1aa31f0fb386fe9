"""numpy restatement of the far-field network (test infrastructure): offset
features, the 4-layer SiLU MLP, the physical envelope, the stress
contraction loss, and its gradient by back-propagation.

``bf=True`` rounds to bfloat16 (round to nearest even) where the HIP path
stores bf16 operands: the features, the weights, the hidden activations, dY
and the activation gradients dZ.  ``bf=False`` is the exact FP64 network
the finite-difference check pins.
"""
import numpy as np

N_IN, N_H, N_OUT, N_CHAN = 128, 512, 5408, 5400
MONO_DEG = 7
N_MOM, N_POLY = 90, 10


def layout():
    """(name, offset, shape) of W1, b1, ..., W4, b4 (row-major, in x out)."""
    shapes = [("W1", (N_IN, N_H)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)),
              ("W3", (N_H, N_H)), ("b3", (N_H,)), ("W4", (N_H, N_OUT)), ("b4", (N_OUT,))]
    out, o = [], 0
    for n, s in shapes:
        out.append((n, o, s))
        o += int(np.prod(s))
    return out, o


def unpack(p):
    lay, _ = layout()
    return {n: np.asarray(p[o:o + int(np.prod(s))], dtype=np.float64).reshape(s)
            for n, o, s in lay}


def pack(d):
    lay, tot = layout()
    p = np.zeros(tot)
    for n, o, s in lay:
        p[o:o + int(np.prod(s))] = d[n].ravel()
    return p


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def features(off, aH=0.0):
    """1/R, 1/R^2, 1/R^3, then direction monomials ux^a uy^b uz^c of degree
    d = 1..7 in the order a = d..0, b = d-a..0, then 1/R_a, 1/R_a^2, 1/R_a^3
    with R_a^2 = R^2 + aH^2; zero padded to N_IN."""
    off = np.asarray(off, dtype=np.float64)
    R = np.sqrt((off ** 2).sum(1))
    u = off / R[:, None]
    f = np.zeros((len(off), N_IN))
    f[:, 0] = 1 / R
    f[:, 1] = 1 / R ** 2
    f[:, 2] = 1 / R ** 3
    q = 3
    for d in range(1, MONO_DEG + 1):
        for a in range(d, -1, -1):
            for b in range(d - a, -1, -1):
                f[:, q] = u[:, 0] ** a * u[:, 1] ** b * u[:, 2] ** (d - a - b)
                q += 1
    ra = 1 / np.sqrt(R ** 2 + aH ** 2)
    f[:, q], f[:, q + 1], f[:, q + 2] = ra, ra ** 2, ra ** 3
    return f


def _orders():
    k = np.arange(N_MOM)
    ok = np.where(k < 9, 0, np.where(k < 36, 1, 2))
    l = np.arange(N_POLY)
    ol = np.where(l == 0, 0, np.where(l < 4, 1, 2))
    return ok, ol


def envelope(off, envScale):
    """env[s, q * 90 + k] = envScale |o_s|^-(2 + order(k) + order(l)), q = c * 10 + l."""
    R = np.sqrt((np.asarray(off, dtype=np.float64) ** 2).sum(1))
    ok, ol = _orders()
    pw = 2 + ol[None, :, None] + ok[None, None, :]                 # (1, 10, 90)
    e = envScale * (1 / R)[:, None, None] ** pw                     # (ns, 10, 90)
    return np.tile(e[:, None], (1, 6, 1, 1)).reshape(len(off), N_CHAN)


def silu(z):
    return z / (1 + np.exp(-z))


def silu_grad(z):
    s = 1 / (1 + np.exp(-z))
    return s * (1 + z * (1 - s))


def forward(p, off, bf=True):
    r = bf16 if bf else (lambda a: a)
    w = unpack(p)
    X = r(features(off))
    cache = {"X": X}
    H = X
    for i in (1, 2, 3):
        Z = H @ r(w[f"W{i}"]) + w[f"b{i}"]
        if bf:
            Z = Z.astype(np.float32).astype(np.float64)
        cache[f"Z{i}"] = Z
        H = r(silu(Z))
        cache[f"H{i}"] = H
    Y = H @ r(w["W4"]) + w["b4"]
    cache["Y"] = Y
    return Y, cache


def loss_dY(Y, off, M, Lt, den, envScale):
    """Predicted local expansions L_hat[s, q] = sum_k env W y M (q = 10 c + l),
    loss sum_s |L_hat_s - Lt_s|^2 / (den_s ns) and its gradient in Y."""
    ns = len(off)
    C = envelope(off, envScale).reshape(ns, 60, N_MOM) * M[:, None, :]
    pred = (C * Y[:, :N_CHAN].reshape(ns, 60, N_MOM)).sum(2)
    res = pred - Lt
    loss = ((res ** 2).sum(1) / (den * ns)).sum()
    dY = np.zeros_like(Y)
    dY[:, :N_CHAN] = (2 * res[:, :, None] * C / (den * ns)[:, None, None]).reshape(ns, N_CHAN)
    return loss, dY, pred


def backward(p, cache, dY, bf=True):
    r = bf16 if bf else (lambda a: a)
    w = unpack(p)
    g = {}
    d = r(dY)
    g["W4"] = cache["H3"].T @ d
    g["b4"] = d.sum(0)
    dH = d @ r(w["W4"]).T
    for i in (3, 2, 1):
        dZ = r(dH * silu_grad(cache[f"Z{i}"]))
        Hin = cache[f"H{i - 1}"] if i > 1 else cache["X"]
        g[f"W{i}"] = Hin.T @ dZ
        g[f"b{i}"] = dZ.sum(0)
        if i > 1:
            dH = dZ @ r(w[f"W{i}"]).T
    return pack(g)


def loss_grad(p, off, M, Lt, den, envScale, bf=True):
    Y, cache = forward(p, off, bf)
    loss, dY, _ = loss_dY(Y, off, M, Lt, den, envScale)
    return loss, backward(p, cache, dY, bf)


def init_params(seed):
    """Random parameters of the same scales as the library's initialisation."""
    rng = np.random.default_rng(seed)
    d = {}
    for i, (fi, fo, gain) in enumerate([(N_IN, N_H, 1.7), (N_H, N_H, 1.7), (N_H, N_H, 1.7),
                                        (N_H, N_OUT, 0.05)], start=1):
        d[f"W{i}"] = rng.standard_normal((fi, fo)) * gain / np.sqrt(fi)
        d[f"b{i}"] = rng.standard_normal(fo) * 0.1
    return pack(d)


def random_batch(ns, seed, span=7):
    rng = np.random.default_rng(seed)
    off = rng.integers(-span, span + 1, size=(ns, 3))
    z = np.all(off == 0, axis=1)
    off[z, 0] = 2
    M = rng.standard_normal((ns, N_MOM))
    Lt = rng.standard_normal((ns, 60)) * 0.05
    den = rng.uniform(0.5, 2.0, size=ns)
    return off.astype(np.int32), M, Lt, den


def projection():
    """Nodes (3) and the 10 x 27 L2 projection onto the quadratic basis of
    the 27-point tensor Gauss rule on the centred unit cell."""
    r = np.sqrt(0.6) / 2
    xi = np.array([-r, 0.0, r])
    w = np.array([5, 8, 5]) / 18
    B, W = [], []
    for p in range(27):
        x, y, z = xi[p % 3], xi[(p // 3) % 3], xi[p // 9]
        B.append([1, x, y, z, x * x, y * y, z * z, x * y, x * z, y * z])
        W.append(w[p % 3] * w[(p // 3) % 3] * w[p // 9])
    B, W = np.array(B), np.array(W)
    return xi, np.linalg.solve(B.T @ (W[:, None] * B), B.T * W)
