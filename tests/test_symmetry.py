"""CPU: the cubic symmetry the hybrid far field uses (W(R o) = det(R) T_L(R)
W(o) T_M(R)^T for the 48 signed axis permutations R), pinned piece by piece:
the oracle's non-singular stress field is isotropic (sigma of the rotated
network at R x is det(R) R sigma(x) R^T: a reflection also reverses the
right-handed Burgers circuit, so b alone does not carry the sign of a
mirror image), the cell moments of a rotated network
are T_M(R) of the original cell's moments, and T_L(R) is the rotation of a
quadratic stress expansion, sigma'(x) = R sigma(R^T x) R^T."""
import numpy as np

import hybrid_ref as H
import pyoracle as O
import refnets as R
import symmetry_ref as S
from dddhip.generate import random_loop_network


def _voigt_to_tensor(s):
    return np.array([[s[0], s[3], s[4]], [s[3], s[1], s[5]], [s[4], s[5], s[2]]])


def test_maps_form_the_group():
    ms = S.maps()
    assert len(ms) == 48
    for Rm in ms:
        TM, TL = S.T_moments(Rm), S.T_locals(Rm)
        for T in (TM, TL):
            assert np.array_equal(np.abs(T).sum(0), np.ones(T.shape[0]))
            assert np.array_equal(T @ T.T, np.eye(T.shape[0]))
    # T_M and T_L are representations: T(R1 R2) = T(R1) T(R2)
    rng = np.random.default_rng(0)
    for _ in range(20):
        a, b = rng.integers(48, size=2)
        A, B = S.matrix(ms[a]), S.matrix(ms[b])
        c = next(k for k, m in enumerate(ms) if np.array_equal(S.matrix(m), A @ B))
        assert np.array_equal(S.T_moments(ms[c]), S.T_moments(ms[a]) @ S.T_moments(ms[b]))
        assert np.array_equal(S.T_locals(ms[c]), S.T_locals(ms[a]) @ S.T_locals(ms[b]))


def test_stress_field_is_isotropic():
    dln, mat = R.params()
    net = random_loop_network(3, seed=0x51)
    centre = np.array([1000.0, 1000.0, 1000.0])
    x = np.random.default_rng(1).uniform(0, 2000, (3, 20))
    s0 = O.calc_σTilde(x, dln, mat, net)
    for k in range(48):
        Rm = S.maps()[k]
        A = S.matrix(Rm)
        s1 = O.calc_σTilde(centre[:, None] + A @ (x - centre[:, None]), dln, mat,
                           S.rotate_network(net, Rm, centre))
        for p in range(x.shape[1]):
            a = S.det(Rm) * A @ _voigt_to_tensor(s0[:, p]) @ A.T
            b = _voigt_to_tensor(s1[:, p])
            assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()


def test_cell_moments_rotate_with_T_M():
    net = random_loop_network(24, seed=0x52)
    G, h = 8, 250.0
    centre = np.array([G * h / 2] * 3)
    M = H.moments(net, G, (0.0, 0.0, 0.0), h)
    for k in (1, 9, 22, 40):
        Rm = S.maps()[k]
        A = S.matrix(Rm)
        M2 = H.moments(S.rotate_network(net, Rm, centre), G, (0.0, 0.0, 0.0), h)
        TM = S.T_moments(Rm)
        for c in range(G ** 3):
            if not M[:, c].any():
                continue
            c3 = np.array([c % G, (c // G) % G, c // (G * G)]) + 0.5 - G / 2
            r3 = np.rint(A @ c3 + G / 2 - 0.5).astype(int)
            c2 = r3[0] + G * (r3[1] + G * r3[2])
            assert np.allclose(M2[:, c2], TM @ M[:, c], rtol=1e-12, atol=1e-9)


def test_local_expansion_rotates_with_T_L():
    rng = np.random.default_rng(3)
    for k in (2, 17, 29, 46):
        Rm = S.maps()[k]
        A = S.matrix(Rm)
        TL = S.T_locals(Rm)
        coef = rng.standard_normal(60)
        coef2 = TL @ coef
        for _ in range(5):
            x = rng.uniform(-0.5, 0.5, 3)
            s = coef.reshape(6, 10) @ H.basis(A.T @ x)
            want = A @ _voigt_to_tensor(s) @ A.T
            got = _voigt_to_tensor(coef2.reshape(6, 10) @ H.basis(x))
            assert np.allclose(got, want, rtol=0, atol=1e-12)
