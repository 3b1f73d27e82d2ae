"""Device transcendentals against the host C library, bit for bit: erf, erfc,
exp and sqrt (ctypes libm), and pow for the rmod operator (libm's pow through
oracle/libpowcheck.so)."""
import ctypes
import ctypes.util

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
    """The rmod operator's pow (engine/sim_core.hpp:35 calls the C
    library's) on the device, bit for bit equal to the host's libm pow on
    7.2 million arguments of rmod's domain and every special case.  The
    misroundings glibc's pow makes against the correctly rounded power are
    shared (tests/test_pow_glibc.py shows them on the host build)."""
    from oracle import PowCheck
    from pow_cases import rmod_args, special_args
    from sigp.api import device_math
    pc = PowCheck()
    for seed in (1, 2, 3):
        x, y = rmod_args(400_000, seed=seed)
        _, lm, bad_host = pc.sweep(x, y)
        assert bad_host == 0
        got = device_math("pow", x, y)
        same = (bits(got) == bits(lm)) | (np.isnan(got) & np.isnan(lm))
        assert same.all(), "%d of %d differ, first x=%r y=%r" % (
            (~same).sum(), same.size, x[~same][0], y[~same][0])
    x, y = special_args()
    _, lm, _ = pc.sweep(x, y)
    got = device_math("pow", x, y)
    same = (bits(got) == bits(lm)) | (np.isnan(got) & np.isnan(lm))
    assert same.all(), list(zip(x[~same], y[~same], got[~same], lm[~same]))[:5]
