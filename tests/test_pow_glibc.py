"""The rmod operator's pow restated from glibc's method (csrc/sigp_pow.h),
compiled for the host, against the host C library's pow -- the function the
reference calls (engine/sim_core.hpp:35) -- bit for bit (CPU only; the
device build of the same header is compared in tests/test_gpu_math.py)."""
import numpy as np

from golden_data import bits
from oracle import PowCheck
from pow_cases import rmod_args, special_args


def test_pow_bitwise_to_libm_rmod_range():
    x, y = rmod_args(400_000, seed=1)            # 2.4 M arguments
    _, _, bad = PowCheck().sweep(x, y)
    assert bad == 0


def test_pow_bitwise_to_libm_special_cases():
    x, y = special_args()
    mine, ref, bad = PowCheck().sweep(x, y)
    assert bad == 0, [(a, b, m, r) for a, b, m, r in zip(x, y, mine, ref)
                      if bits(m) != bits(r) and not (np.isnan(m) and np.isnan(r))][:5]


def _correctly_rounded_pow(x: float, y: float) -> float:
    """x^y to 60 digits, rounded once to double (x > 0 normal, result normal)."""
    from decimal import Decimal, getcontext
    getcontext().prec = 60
    return float((Decimal(x).ln() * Decimal(y)).exp())


def test_pow_shares_libm_misroundings():
    """glibc's pow is not correctly rounded.  Among rmod arguments whose
    correctly rounded power (60-digit decimals) differs from libm's, the
    restatement gives libm's value, not the correctly rounded one."""
    rng = np.random.default_rng(9)
    x = np.exp2(rng.random(60_000) * 30 - 15)
    y = 1.0 + rng.random(60_000)
    mine, ref, bad = PowCheck().sweep(x, y)
    assert bad == 0
    cr = np.array([_correctly_rounded_pow(a, b) for a, b in zip(x, y)])
    mis = np.flatnonzero(cr != ref)
    assert mis.size >= 5                          # libm misrounds a few in 10^4
    assert np.array_equal(bits(mine[mis]), bits(ref[mis]))
    assert np.all(np.abs(cr[mis] - ref[mis]) <= np.spacing(np.abs(ref[mis])))
