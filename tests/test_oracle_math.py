"""The double-precision transcendentals the oracle and the GPU share (own
routines: Cody-Waite + Taylor sin/cos, reduced Taylor atan, atan2/asin/acos
on top) against the C library: within a few ulp in double, and identical
after rounding to float -- which is how the path tracer uses them -- for all
but a vanishing fraction of arguments; zeros, signs and special values as
the C library gives them."""
import math

import numpy as np
import pytest

LIB = {"sin": math.sin, "cos": math.cos, "atan": math.atan, "asin": math.asin, "acos": math.acos}


def _lib(fn, x, y=None):
    if fn == "atan2":
        return np.array([math.atan2(a, b) for a, b in zip(y, x)])
    f = LIB[fn]
    return np.array([f(t) for t in x])


def _args(fn, rng):
    if fn in ("sin", "cos"):
        return np.concatenate([rng.uniform(-7, 7, 200000), rng.uniform(-3000, 3000, 50000),
                               np.arange(-64, 65) * (np.pi / 4),
                               np.float64(np.float32(rng.uniform(0, 2 * np.pi, 50000)))])
    if fn == "atan":
        return np.concatenate([rng.uniform(-2, 2, 100000), np.tan(rng.uniform(-1.57, 1.57, 100000)),
                               rng.standard_cauchy(50000), [0.2679491924311227, 1.0, -1.0, 1e-300, 1e300]])
    if fn in ("asin", "acos"):
        return np.concatenate([rng.uniform(-1, 1, 200000), 1 - rng.uniform(0, 1e-6, 20000),
                               -1 + rng.uniform(0, 1e-6, 20000), [-1.0, 1.0, 0.5, -0.5, 1e-20]])
    raise AssertionError(fn)


@pytest.mark.parametrize("fn,ulps", [("sin", 2), ("cos", 2), ("atan", 3), ("asin", 4), ("acos", 4)])
def test_within_ulps_of_library_and_equal_as_float(oracle, fn, ulps):
    rng = np.random.default_rng(hash(fn) % 1000)
    x = _args(fn, rng)
    got, ref = oracle.math_d(fn, x), _lib(fn, x)
    assert np.all(np.abs(got - ref) <= ulps * np.spacing(np.abs(ref)) + 1e-300), fn
    diff = np.count_nonzero(got.astype(np.float32) != ref.astype(np.float32))
    assert diff <= len(x) * 1e-5, (fn, diff)


def test_atan2_quadrants_zeros_and_ulps(oracle):
    rng = np.random.default_rng(11)
    d = rng.normal(size=(200000, 2))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    y, x = d[:, 0], d[:, 1]
    got, ref = oracle.math_d("atan2", x, y), _lib("atan2", x, y)
    assert np.all(np.abs(got - ref) <= 4 * np.spacing(np.abs(ref)))
    assert np.count_nonzero(got.astype(np.float32) != ref.astype(np.float32)) <= 2
    # signed zeros, axes and infinities follow the C library exactly
    zs = [0.0, -0.0, 1.0, -1.0, np.inf, -np.inf]
    ys, xs = np.array([a for a in zs for b in zs]), np.array([b for a in zs for b in zs])
    got, ref = oracle.math_d("atan2", xs, ys), _lib("atan2", xs, ys)
    assert np.array_equal(got, ref) and np.array_equal(np.signbit(got), np.signbit(ref))


def test_special_values(oracle):
    s = oracle.math_d("sin", np.array([0.0, -0.0, np.pi / 2, 2.0 ** 19, np.inf, np.nan]))
    assert s[0] == 0 and not np.signbit(s[0]) and s[1] == 0 and np.signbit(s[1]) and s[2] == 1.0
    assert s[3] == math.sin(2.0 ** 19) and np.isnan(s[4]) and np.isnan(s[5])
    a = oracle.math_d("atan", np.array([0.0, -0.0, np.inf, -np.inf, np.nan]))
    assert a[0] == 0 and np.signbit(a[1]) and a[2] == math.pi / 2 and a[3] == -math.pi / 2 and np.isnan(a[4])
    assert np.all(np.isnan(oracle.math_d("asin", np.array([1.5, -2.0, np.nan]))))
    assert oracle.math_d("acos", np.array([-1.0]))[0] == math.pi
    y = np.random.default_rng(1).uniform(-50, 50, 100000)
    s, c = oracle.math_d("sin", y), oracle.math_d("cos", y)
    assert np.max(np.abs(s * s + c * c - 1.0)) < 4e-16
