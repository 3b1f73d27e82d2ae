"""Helpers shared by the GPU parity tests."""
import numpy as np

from golden_data import bits


def assert_bitwise(got, want, what=""):
    g, w = bits(got), bits(want)
    if not np.array_equal(g, w):
        idx = np.flatnonzero(g != w)
        k = idx[0]
        raise AssertionError("%s: %d of %d differ; first at %d: got %r want %r"
                             % (what, idx.size, g.size, k, np.asarray(got).ravel()[k],
                                np.asarray(want).ravel()[k]))
