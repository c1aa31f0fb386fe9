"""CPU: pins of the numpy far-field network restatement.

The back-propagated gradient of the exact (FP64, no bf16) network is checked
against central finite differences of the loss, one parameter block at a
time, so that a dropped term or a transposed operand in any layer fails; the
features and envelope are checked against hand-computed values."""
import numpy as np

import mlp_ref as R


def test_features_known_offset():
    f = R.features(np.array([[3, 0, 4]]))[0]
    assert np.allclose(f[:3], [0.2, 0.04, 0.008])
    # degree 1: ux, uy, uz; degree 2: ux^2, ux uy, ux uz, uy^2, uy uz, uz^2
    assert np.allclose(f[3:12], [0.6, 0, 0.8, 0.36, 0, 0.48, 0, 0, 0.64])
    # last monomial uz^7, then 1/R_a powers (aH = 0: 1/R), then padding
    assert np.isclose(f[3 + 119 - 1], 0.8 ** 7)
    assert np.allclose(f[122:125], [0.2, 0.04, 0.008])
    assert not f[125:].any()
    g = R.features(np.array([[3, 0, 4]]), aH=np.sqrt(11.0))[0]  # R_a = 6
    assert np.allclose(g[122:125], [1 / 6, 1 / 36, 1 / 216])


def test_envelope_powers():
    off = np.array([[0, 2, 0]])
    e = R.envelope(off, 3.0).reshape(6, 10, 90)
    # l = 0, k = 0: |o|^-2; l = 5 (quadratic), k = 40 (second order): |o|^-6
    assert np.isclose(e[2, 0, 0], 3.0 / 4) and np.isclose(e[4, 5, 40], 3.0 / 64)
    assert np.isclose(e[0, 2, 20], 3.0 / 16)


def test_contraction_direct_sum():
    off, M, Lt, den = R.random_batch(3, seed=5)
    Y = np.random.default_rng(6).standard_normal((3, R.N_OUT))
    loss, dY, pred = R.loss_dY(Y, off, M, Lt, den, 2.0)
    ok, ol = R._orders()
    s = 1
    Rn = np.linalg.norm(off[s])
    want = np.zeros(60)
    for c in range(6):
        for l in range(10):
            for k in range(90):
                q = c * 10 + l
                want[q] += 2.0 * Rn ** -(2 + ok[k] + ol[l]) * Y[s, q * 90 + k] * M[s, k]
    assert np.allclose(pred[s], want, rtol=1e-12)
    assert np.isclose(loss, sum(((pred[i] - Lt[i]) ** 2).sum() / (den[i] * 3) for i in range(3)))


def test_projection_reproduces_quadratics():
    """The 27-point projection returns the coefficients of any quadratic and
    the L2-best quadratic of a cubic (its cubic part is orthogonal to x^3 -
    (3/20) x on the centred unit interval)."""
    xi, Pj = R.projection()
    pts = np.array([[xi[p % 3], xi[(p // 3) % 3], xi[p // 9]] for p in range(27)])
    x, y, z = pts.T
    c = np.arange(1, 11, dtype=float)
    f = c @ np.stack([np.ones(27), x, y, z, x * x, y * y, z * z, x * y, x * z, y * z])
    assert np.allclose(Pj @ f, c)
    # x^3 = (x^3 - 3x/20) + 3x/20: the projection keeps only 3/20 x
    assert np.allclose(Pj @ x ** 3, [0, 0.15, 0, 0, 0, 0, 0, 0, 0, 0])


def test_backprop_matches_finite_differences():
    p = R.init_params(1)
    off, M, Lt, den = R.random_batch(12, seed=2)
    es = 0.7
    _, g = R.loss_grad(p, off, M, Lt, den, es, bf=False)
    lay, _ = R.layout()
    rng = np.random.default_rng(3)
    for name, o, shp in lay:
        n = int(np.prod(shp))
        v = np.zeros_like(p)
        v[o:o + n] = rng.standard_normal(n)
        eps = 1e-6
        lp = R.loss_grad(p + eps * v, off, M, Lt, den, es, bf=False)[0]
        lm = R.loss_grad(p - eps * v, off, M, Lt, den, es, bf=False)[0]
        fd = (lp - lm) / (2 * eps)
        an = g @ v
        assert abs(fd - an) <= 1e-6 * abs(an) + 1e-12, (name, fd, an)


def test_bf16_rounding_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5], dtype=np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1) wins
    assert list(R.bf16(x)) == [1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5]
