"""numpy restatement of the far-field network (test infrastructure): offset
features, the 4-layer SiLU MLP, the physical envelope, the stress
contraction loss, and its gradient by back-propagation.

``bf=True`` rounds to bfloat16 (round to nearest even) where the HIP path
stores bf16 operands: the features, the weights, the hidden activations, dY
and the activation gradients dZ.  ``bf=False`` is the exact FP64 network
the finite-difference check pins.
"""
import numpy as np

N_IN, N_H, N_OUT, N_CHAN = 32, 256, 5408, 5400
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


def features(off):
    off = np.asarray(off, dtype=np.float64)
    R = np.sqrt((off ** 2).sum(1))
    u = off / R[:, None]
    f = np.zeros((len(off), N_IN))
    f[:, 0:3] = u
    f[:, 3] = 1 / R
    f[:, 4] = 1 / R ** 2
    f[:, 5:8] = u ** 2
    f[:, 8] = u[:, 0] * u[:, 1]
    f[:, 9] = u[:, 0] * u[:, 2]
    f[:, 10] = u[:, 1] * u[:, 2]
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


def loss_dY(Y, off, M, P, sig, den, envScale):
    ns = len(off)
    C = envelope(off, envScale).reshape(ns, 6, N_POLY, N_MOM)
    C = C * P[:, None, :, None] * M[:, None, None, :]
    pred = (C * Y[:, :N_CHAN].reshape(ns, 6, N_POLY, N_MOM)).sum((2, 3))
    res = pred - sig
    loss = ((res ** 2).sum(1) / (den * ns)).sum()
    dY = np.zeros_like(Y)
    dY[:, :N_CHAN] = (2 * res[:, :, None, None] * C / (den * ns)[:, None, None, None]).reshape(
        ns, N_CHAN)
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


def loss_grad(p, off, M, P, sig, den, envScale, bf=True):
    Y, cache = forward(p, off, bf)
    loss, dY, _ = loss_dY(Y, off, M, P, sig, den, envScale)
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
    rho = rng.uniform(-0.5, 0.5, size=(ns, 3))
    x, y, zz = rho.T
    P = np.stack([np.ones(ns), x, y, zz, x * x, y * y, zz * zz, x * y, x * zz, y * zz], 1)
    sig = rng.standard_normal((ns, 6)) * 0.05
    den = rng.uniform(0.5, 2.0, size=ns)
    return off.astype(np.int32), M, P, sig, den
