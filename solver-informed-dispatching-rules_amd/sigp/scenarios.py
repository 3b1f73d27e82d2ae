"""Stochastic service-time scenarios with common random numbers.

Gaussian mode restates engine/sim_host.hpp:42-58: scenario s draws from a
fresh generator seeded base_seed + s; customer i (in index order) gets
max(0, gauss(mu_i, sigma * mu_i)) and customers with mu_i <= 0 draw nothing.
The uniform stream is CPython's Mersenne Twister, which engine/rng.hpp:13-93
restates.  The Box-Muller pair (engine/rng.hpp:135-147) is computed as the
reference's optimised build computes it: cos and sin of the same angle in
one libm sincos() call (g++ -O2 fuses the two calls, and glibc's sincos
rounds differently from separate sin/cos in ~0.07% of arguments), so the
rows are bit-identical to the reference's gen_durations on the same libm.

Gamma mode is new (the reference has no Gamma sampler): customer i gets
Gamma(shape k, scale mu_i / k) -- mean mu_i -- drawn by Marsaglia & Tsang's
squeeze method (ACM TOMS 26(3), 2000) from the same per-scenario stream
(gauss for the normal variate, random() for the uniform); shape < 1 uses
Gamma(k + 1) * U^(1/k).  Customers with mu_i <= 0 again draw nothing.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import math
import random

import numpy as np

_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
_libm.sincos.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                         ctypes.POINTER(ctypes.c_double)]
_libm.sincos.restype = None
_TWOPI = 2.0 * 3.14159265358979323846


class GaussStream:
    """Mt19937 + cached Box-Muller gauss (engine/rng.hpp:135-147).

    `math`, when given, supplies the log and sincos used instead of libm's
    (an object with .log(x) and .sincos(x) -> (sin, cos)); the device sampler
    is this stream with correctly rounded functions."""

    __slots__ = ("rng", "has", "spare", "_s", "_c", "log", "_math")

    def __init__(self, seed: int, math_fns=None):
        self.rng = random.Random(seed)
        self.has = False
        self.spare = 0.0
        self._s = ctypes.c_double()
        self._c = ctypes.c_double()
        self._math = math_fns
        self.log = math.log if math_fns is None else math_fns.log

    def random(self) -> float:
        return self.rng.random()

    def gauss(self, mu: float, sigma: float) -> float:
        if self.has:
            z = self.spare
            self.has = False
        else:
            x2pi = self.rng.random() * _TWOPI
            g2rad = math.sqrt(-2.0 * self.log(1.0 - self.rng.random()))
            if self._math is None:
                _libm.sincos(x2pi, ctypes.byref(self._s), ctypes.byref(self._c))
                sn, cs = self._s.value, self._c.value
            else:
                sn, cs = self._math.sincos(x2pi)
            z = cs * g2rad
            self.spare = sn * g2rad
            self.has = True
        return mu + z * sigma

__all__ = ["gen_durations", "gamma_draw", "GaussStream"]


def gamma_draw(rng: GaussStream, shape: float, scale: float) -> float:
    if shape < 1.0:
        g = gamma_draw(rng, shape + 1.0, 1.0)
        u = rng.random()
        return scale * g * (u ** (1.0 / shape))
    d = shape - 1.0 / 3.0
    c = 1.0 / math.sqrt(9.0 * d)
    while True:
        x = rng.gauss(0.0, 1.0)
        v = 1.0 + c * x
        if v <= 0.0:
            continue
        v = v * v * v
        u = rng.random()
        if u < 1.0 - 0.0331 * (x * x) * (x * x):
            return scale * d * v
        if u > 0.0 and rng.log(u) < 0.5 * x * x + d * (1.0 - v + rng.log(v)):
            return scale * d * v


def gen_durations(expected: np.ndarray, n_samples: int, base_seed: int, sigma: float = 0.2,
                  mode: str = "gauss", shape: float = 4.0, math_fns=None) -> np.ndarray:
    """n_samples x n row-major array (scenario-major), one generator per
    scenario; `math_fns` replaces libm's log and sincos (see GaussStream)."""
    mu = [float(v) for v in np.asarray(expected, dtype=np.float64)]
    n = len(mu)
    out = np.zeros((n_samples, n), dtype=np.float64)
    for s in range(n_samples):
        rng = GaussStream(base_seed + s, math_fns)
        row = out[s]
        for i, e in enumerate(mu):
            if e <= 0.0:
                row[i] = 0.0
                continue
            if mode == "gauss":
                v = rng.gauss(e, sigma * e)
                row[i] = v if v > 0.0 else 0.0
            elif mode == "gamma":
                row[i] = gamma_draw(rng, shape, e / shape)
            else:
                raise ValueError("mode must be 'gauss' or 'gamma'")
    return np.ascontiguousarray(out)
