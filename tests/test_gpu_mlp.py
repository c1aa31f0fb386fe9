"""GPU: the far-field network.

* loss and gradient of ddd_mlp_loss_grad (BF16 MFMA products, FP32
  accumulation) against the numpy restatement that rounds to bf16 at the same
  points: loss within 1e-4, each parameter block's gradient within 3e-3
  normwise (FP32 vs FP64 accumulation and bf16 ties that fall differently).
  At 4,096 samples (the training batch) the output layer's three products
  run on the 256 x 256 tiles (one with split K over a ragged K), the
  smaller batches on the 128 x 128 ones; at 4,100 samples the 256 x 256
  tiles have a ragged last row tile, and at 4,100 and 99 samples the weight
  gradients (K not a multiple of 8) take the generic kernel;
* training on the exact stresses of one network lowers the loss, and the
  operator it predicts improves the hybrid forces of another network over the
  no-far-field ablation and over the untrained network.
"""
import numpy as np
import pytest

import mlp_ref as R
import pyoracle as O
import refnets as RN
from dddhip import forces as G
from dddhip.generate import random_loop_network
from metrics import normwise

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def prm():
    return RN.params()


def test_params_roundtrip():
    m = G.FarFieldMLP(seed=7)
    p = m.get_params()
    assert p.size == R.layout()[1] and np.isfinite(p).all()
    q = R.init_params(9).astype(np.float32)
    m.set_params(q)
    assert np.array_equal(m.get_params(), q)
    m.close()


@pytest.mark.parametrize("ns", [4100, 4096, 1000, 200, 99, 64])
def test_loss_grad_vs_reference(ns):
    m = G.FarFieldMLP(seed=1)
    p = R.init_params(3).astype(np.float32)
    m.set_params(p)
    off, M, Lt, den = R.random_batch(ns, seed=ns)
    f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)
    M, Lt, den = f32(M), f32(Lt), f32(den)
    es = 0.5
    loss, g = m.loss_grad(off, M, Lt, den, es)
    lr, gr = R.loss_grad(p.astype(np.float64), off, M, Lt, den, es, bf=True)
    print(f"ns={ns}: loss {loss:.6e} ref {lr:.6e}", flush=True)
    assert abs(loss - lr) <= 1e-4 * abs(lr)
    for name, o, shp in R.layout()[0]:
        n = int(np.prod(shp))
        e = normwise(g[o:o + n], gr[o:o + n])
        print(f"  {name}: normwise {e:.2e}", flush=True)
        assert e <= 3e-3, name
    m.close()


def test_local_targets_vs_oracle(prm):
    """Training targets: the oracle's field stress (calc_σTilde restricted to
    the source cell's segments) at the 27 Gauss points of the target cell,
    projected onto the quadratic basis; FP32 output, 1e-6 relative."""
    import hybrid_ref as H
    dln, mat = prm
    Gc, h = 8, 250.0
    net = random_loop_network(64, seed=0xE3)
    ff = G.FarField(Gc, (0.0, 0.0, 0.0), h, 1)
    cell = H.bin_cells(net, Gc, (0, 0, 0), h)
    rng = np.random.default_rng(1)
    live = np.unique(cell[cell >= 0])
    src = rng.choice(live, 6)
    off = np.array([[2, 0, 0], [-3, 1, 2], [0, 0, -2], [5, -4, 1], [2, 2, 2], [-7, 7, 0]])
    got = G.farFieldLocalTargets(dln, mat, ff, net, src, off)
    xi, Pj = R.projection()
    for s in range(len(src)):
        c = src[s]
        tc = (np.array([c % Gc, (c // Gc) % Gc, c // (Gc * Gc)]) - off[s] + 0.5) * h
        pts = np.array([tc + h * np.array([xi[p % 3], xi[(p // 3) % 3], xi[p // 9]])
                        for p in range(27)])
        sg = O.calc_σTilde(pts.T, dln, mat, net, idx=np.nonzero(cell == c)[0] + 1)
        want = (sg @ Pj.T).ravel()  # 6 x 10, component-major
        err = np.abs(got[s] - want).max() / np.abs(want).max()
        print(f"sample {s}: cell {c} offset {off[s]}: rel {err:.2e}", flush=True)
        assert err <= 1e-6
    ff.close()


def test_training_and_ablation(prm):
    dln, mat = prm
    Gc, rn, h = 8, 1, 250.0
    train = random_loop_network(512, seed=0xE1)
    ev = random_loop_network(192, seed=0xE2)
    ff = G.FarField(Gc, (0.0, 0.0, 0.0), h, rn)
    S = 2 * Gc - 1
    ref = O.calcSegSegForcePar(dln, mat, ev, 16)

    def err():
        g, _, _ = G.calcSegSegForceHybrid(dln, mat, ev, ff, farDist=1e30)
        return normwise(g, ref)

    ff.set_operator(np.zeros((S ** 3, 60, 90), np.float32))
    e_none = err()
    m = G.FarFieldMLP(seed=3)
    G.setFarFieldFromMLP(dln, mat, ff, m)
    e_untrained = err()
    rep = G.trainFarField(dln, mat, ff, m, train, samples=1 << 17, valSamples=1 << 13, steps=600,
                          batch=1024, lr=1e-3, lrFinal=1e-4, seed=5)
    e_trained = err()
    print(f"report {rep}", flush=True)
    print(f"seg-seg normwise error: no far field {e_none:.3e}, untrained {e_untrained:.3e}, "
          f"trained {e_trained:.3e}", flush=True)
    assert rep["lossLast"] < 0.3 * rep["lossFirst"]
    # held-out error of all 60 local-expansion coefficients, falling with distance
    assert rep["relErrVal"] < 0.8 and rep["relErrShell"][5] < rep["relErrShell"][0]
    assert e_trained < 0.7 * e_none and e_trained < 0.7 * e_untrained
    m.close()
    ff.close()


def _cell_map(Rm, G):
    """Cell index t -> the cell R t of the grid rotated about its centre."""
    import symmetry_ref as S
    A = S.matrix(Rm)
    t = np.arange(G ** 3)
    c3 = np.stack([t % G, (t // G) % G, t // (G * G)]).astype(float) + 0.5 - G / 2
    r3 = np.rint(A @ c3 + G / 2 - 0.5).astype(int)
    return r3[0] + G * (r3[1] + G * r3[2])


@pytest.mark.parametrize("cubic", [True, False])
def test_operator_cubic_equivariance(prm, cubic):
    """The installed far-field operator of a network that learns W at the
    canonical offsets is exactly equivariant: the local expansions of a
    network rotated or reflected about the grid centre are
    det(R) T_L(R) of the original cells' expansions, up to the FP32 FFT
    (an untrained network, so nothing but the symmetry makes it so).  A
    network that learns every offset on its own is not."""
    import symmetry_ref as S
    dln, mat = prm
    Gc, rn, h = 8, 1, 250.0
    net = random_loop_network(24, seed=0x5B)
    ff = G.FarField(Gc, (0.0, 0.0, 0.0), h, rn)
    m = G.FarFieldMLP(seed=4)
    if not cubic:  # ten steps, only to switch the network to every offset
        G.trainFarField(dln, mat, ff, m, net, samples=1 << 12, valSamples=256, steps=10,
                        batch=256, lr=1e-4, lrFinal=1e-4, seed=1, cubic=False)
    G.setFarFieldFromMLP(dln, mat, ff, m)
    G.calcSegSegForceHybrid(dln, mat, net, ff, farDist=1e30)
    L0 = ff.debug(1).astype(np.float64)
    worst = 0.0
    for k in (9, 21, 38, 47):
        Rm = S.maps()[k]
        G.calcSegSegForceHybrid(dln, mat, S.rotate_network(net, Rm, (1000.0,) * 3), ff,
                                farDist=1e30)
        L1 = ff.debug(1).astype(np.float64)
        want = S.det(Rm) * S.T_locals(Rm) @ L0
        got = L1[:, _cell_map(Rm, Gc)]
        worst = max(worst, np.linalg.norm(got - want) / np.linalg.norm(want))
    print(f"cubic={cubic}: worst relative mismatch {worst:.2e}", flush=True)
    if cubic:
        assert worst <= 1e-4
    else:
        assert worst >= 1e-2
    m.close()
    ff.close()
