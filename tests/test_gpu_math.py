"""Device transcendentals against the host C library (ctypes libm), bit for
bit for erf, erfc, exp and sqrt; pow (rmod only) against a 100-digit decimal
reference and against libm."""
import ctypes
import ctypes.util
from decimal import Decimal, getcontext

import numpy as np
import pytest

from golden_data import bits

pytestmark = pytest.mark.gpu
LIBM = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
for _f in ("erf", "erfc", "exp", "sqrt"):
    getattr(LIBM, _f).argtypes = [ctypes.c_double]
    getattr(LIBM, _f).restype = ctypes.c_double
LIBM.pow.argtypes = [ctypes.c_double, ctypes.c_double]
LIBM.pow.restype = ctypes.c_double


def host(fn, xs):
    f = getattr(LIBM, fn)
    return np.array([f(float(v)) for v in xs])


def sample_args(rng, n):
    parts = [rng.uniform(-7, 7, n), rng.uniform(-0.9, 0.9, n // 4), rng.uniform(-30, 30, n // 4),
             rng.normal(0, 3, n // 4), np.array([0.0, -0.0, 1e-300, -1e-300, 6.0, -6.0, 27.9, 28.0,
                                                 0.84375, 1.25, 2.857142857142857, np.inf, -np.inf])]
    return np.concatenate(parts)


@pytest.mark.parametrize("fn", ["erf", "erfc"])
def test_erf_family_bitwise(fn):
    from sigp.api import device_math
    xs = sample_args(np.random.default_rng(7), 200_000)
    assert np.array_equal(bits(device_math(fn, xs)), bits(host(fn, xs)))


def test_exp_bitwise():
    from sigp.api import device_math
    rng = np.random.default_rng(8)
    xs = np.concatenate([rng.uniform(-745, 709, 200_000), rng.uniform(-40, 1, 100_000),
                         np.array([0.0, -0.0, 1e-20, 709.7, -744.9, 800.0, -800.0])])
    assert np.array_equal(bits(device_math("exp", xs)), bits(host("exp", xs)))


def test_sqrt_correctly_rounded():
    from sigp.api import device_math
    rng = np.random.default_rng(9)
    xs = np.concatenate([rng.uniform(0, 1e4, 100_000), np.exp(rng.uniform(-700, 700, 100_000))])
    assert np.array_equal(bits(device_math("sqrt", xs)), bits(np.sqrt(xs)))


def test_pow_for_rmod():
    from sigp.api import device_math
    getcontext().prec = 60
    rng = np.random.default_rng(10)
    x = np.concatenate([rng.uniform(0, 50, 4000), np.exp(rng.uniform(-30, 30, 2000))])
    y = 1.0 + np.concatenate([rng.uniform(0, 1, 5000), np.array([0.0, 1.0] * 500)])
    got = device_math("pow", x, y)
    exact = np.array([float(Decimal(float(a)) ** Decimal(float(b))) for a, b in zip(x, y)])
    assert np.array_equal(bits(got), bits(exact)), "not correctly rounded"
    lm = np.array([LIBM.pow(float(a), float(b)) for a, b in zip(x, y)])
    agree = np.mean(bits(got) == bits(lm))
    print("device pow == libm pow on %.4f of calls" % agree)
    assert agree > 0.995
