"""Arguments for the rmod pow sweeps (CPU and GPU tests share them).

rmod evaluates pow(max(a, 0), 1 + clip(b, 0, 1)) (engine/sim_core.hpp:29-35):
x >= 0 (or NaN), y in [1, 2] with y = 1 and y = 2 common (clipped b).  The
mix covers the magnitudes priorities take (terminal values and their
products, 2^-20 .. 2^20), every positive double, results in the subnormal
and overflow ranges, x near 1 (where log x is tiny and cancellation is
worst), the clip ends, and the three-decimal constants rules carry.
"""
import numpy as np


def rmod_args(n_per_mode: int, seed: int = 0):
    rng = np.random.default_rng(seed)
    u = lambda: rng.random(n_per_mode)
    xs, ys = [], []
    # typical priorities
    xs.append(np.exp2(u() * 40 - 20)); ys.append(1.0 + u())
    # any positive finite double
    bits = rng.integers(0, 0x7FEFFFFFFFFFFFFF, n_per_mode, dtype=np.int64, endpoint=True)
    xs.append(bits.view(np.float64)); ys.append(1.0 + u())
    # subnormal x and results in the subnormal / overflow ranges
    xs.append(np.exp2(u() * 2097.9 - 1074)); ys.append(1.0 + u())
    # x near 1
    xs.append(1.0 + (u() - 0.5) * 2.0 ** -6); ys.append(1.0 + u())
    # clip ends y = 1, y = 2
    xs.append(np.exp2(u() * 60 - 30)); ys.append(np.where(u() < 0.5, 1.0, 2.0))
    # three-decimal values, as constants and sums of them
    xs.append(rng.integers(0, 100000, n_per_mode) / 1000.0)
    ys.append(1.0 + rng.integers(0, 1001, n_per_mode) / 1000.0)
    return np.concatenate(xs), np.concatenate(ys)


def special_args():
    """Every special-case branch of pow, rmod's domain and beyond."""
    xv = [0.0, -0.0, 5e-324, 2.2250738585072009e-308, 2.2250738585072014e-308, 1e-300, 0.5, 1.0,
          1.0000000000000002, 0.9999999999999999, 2.0, 10.0, 1e300, 1.7976931348623157e308,
          float("inf"), float("nan"), -1.0, -2.0, -0.5, -float("inf"), -1e-310]
    yv = [0.0, -0.0, 1.0, 2.0, 1.5, 1.0000000000000002, 1.9999999999999998, 0.5, -1.0, 3.0, -3.0,
          2.0 ** -70, -(2.0 ** -70), 2.0 ** 70, -(2.0 ** 70), float("inf"), -float("inf"), float("nan")]
    x, y = np.meshgrid(np.array(xv), np.array(yv))
    return x.ravel(), y.ravel()
