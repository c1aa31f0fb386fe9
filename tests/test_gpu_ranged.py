"""The all-pairs tile kernel's division, square root, log and atan without
range scaling (ddd.jl_amd/csrc/fp64_ranged.hpp), run on the GPU through
ddd_selftest_ranged and compared bit for bit with the general IEEE
operations and the oracle's log/atan (ddd_libm, the functions the oracle's
pair formula calls in place of DislocationForces.jl's log/atan,
:1480-1561).  The tile kernel relies on these agreeing on every argument
it admits, and on a zero, infinite or NaN operand giving NaN (its tile is
then evaluated again with the general sequences)."""
import numpy as np
import pytest

import pyoracle as O
from dddhip import forces as G

pytestmark = pytest.mark.gpu

rng = np.random.default_rng(0x5EED)


def _same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    nan = np.isnan(a) | np.isnan(b)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    bad = np.flatnonzero(a[~nan].view(np.int64) != b[~nan].view(np.int64))
    assert bad.size == 0, (bad.size, a[~nan][bad[:5]], b[~nan][bad[:5]])


def _around(v):
    v = np.asarray(v, dtype=np.float64)
    return np.concatenate([v, np.nextafter(v, np.inf), np.nextafter(v, -np.inf)])


def test_atan_bitwise():
    """Every interval of the reduction (bounds 7/16, 11/16, 19/16, 39/16 and
    their neighbours), both signs, the 2^-28 and 2^66 clamps, zeros and
    infinities."""
    x = np.concatenate([
        rng.normal(size=400_000) * 10.0 ** rng.uniform(-12, 12, 400_000),
        rng.uniform(-3.0, 3.0, 400_000),
        rng.uniform(-1e6, 1e6, 100_000),
        _around([0.4375, 0.6875, 1.1875, 2.4375, 3.7252902984e-09, 7.378697629483821e+19,
                 1.0, 0.5, 1.5, 2.0 ** -28, 2.0 ** 66, 1e300, 5e-324, 2.2250738585072014e-308]),
        [0.0, -0.0, np.inf, -np.inf, np.nan]])
    x = np.concatenate([x, -x])
    _same_bits(G.selftest_ranged(1, x), O.libm_vec(1, x))


def test_log_bitwise_on_positive_normals():
    x = np.concatenate([np.exp(rng.uniform(-700, 700, 400_000)), rng.uniform(0.5, 2.0, 400_000),
                        1.0 + rng.normal(size=100_000) * 1e-7,
                        _around([1.0, 2.0, 0.5, np.sqrt(0.5), np.sqrt(2.0), 1e300]),
                        [2.2250738585072014e-308, 1.7976931348623157e308]])
    _same_bits(G.selftest_ranged(0, x), O.libm_vec(0, x))


def test_log_nan_outside_positive_normals():
    x = np.array([0.0, -0.0, -1.0, -1e-300, 5e-324, 1e-310, np.inf, -np.inf, np.nan])
    assert np.all(np.isnan(G.selftest_ranged(0, x)))


def test_sqrt_and_division_bitwise():
    x = np.concatenate([10.0 ** rng.uniform(-230, 300, 400_000), rng.uniform(0.0, 4.0, 400_000)[1:],
                        _around([1.0, 2.0, 4.0, 0.25, 2.0 ** -766, 1e300])])
    _same_bits(G.selftest_ranged(2, x), np.sqrt(x))
    a = rng.normal(size=400_000) * 10.0 ** rng.uniform(-100, 100, 400_000)
    b = rng.normal(size=400_000) * 10.0 ** rng.uniform(-100, 100, 400_000)
    ab = np.stack([a, b], axis=1).ravel()
    _same_bits(G.selftest_ranged(3, ab), a / b)
    zero = G.selftest_ranged(3, np.array([1.0, 0.0, 1.0, np.inf, 0.0, 0.0]))
    assert np.all(np.isnan(zero))
    assert np.all(np.isnan(G.selftest_ranged(2, np.array([0.0, np.inf, -1.0]))))
