"""The surrogate network's CPU oracle: its float64 gradients match central
finite differences, its BF16 rounding matches the definition (round to
nearest even on the float32 bits), and its Adam step gives what the
published update fixes."""
import numpy as np

import mlp_oracle as mo


def _params(rng, scale=0.05):
    return rng.normal(0.0, scale, mo.NP)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(3)
    p = _params(rng)
    x = rng.random((6, mo.IN))
    t = rng.normal(size=6)
    _, g = mo.forward_backward(p, x, t, emulate=False)
    idx = np.concatenate([rng.choice(np.arange(mo.O_W1, mo.O_B1), 12), rng.choice(np.arange(mo.O_B1, mo.O_W2), 4),
                          rng.choice(np.arange(mo.O_W2, mo.O_B2), 12), rng.choice(np.arange(mo.O_B2, mo.O_W3), 4),
                          rng.choice(np.arange(mo.O_W3, mo.O_B3), 6), [mo.O_B3]])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (mo.loss_only(pp, x, t) - mo.loss_only(pm, x, t)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(g[k]), (k, fd, g[k])


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, -2.5, 0.0, 3.0e-3], dtype=np.float32)
    got = mo.bf16(v)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to the even mantissa (1)
    assert got[0] == 1.0 and got[1] == 1.0 and got[3] == 1.0
    assert got[2] == 1.0 + 2 ** -7
    assert got[4] == -2.5 and got[5] == 0.0
    assert abs(got[6] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def test_adam_step_closed_forms():
    """With zero moments the first bias-corrected step is lr * g / (|g| + eps);
    under a constant gradient the bias-corrected moments stay g and g^2, so
    every step moves by the same amount.  A wrong beta, a dropped bias
    correction or a transposed m / v each break one of these."""
    rng = np.random.default_rng(8)
    g = (rng.normal(size=4000) * 10.0 ** rng.uniform(-4, 1, 4000)).astype(np.float32)
    p0 = rng.normal(size=4000).astype(np.float32)
    lr = np.float32(1e-3)
    z = np.zeros_like(g)
    p1, m1, v1 = mo.adam_step(p0, g, z, z, 1, lr)
    want = lr * g.astype(np.float64) / (np.abs(g.astype(np.float64)) + 1e-8)
    assert np.allclose(p0.astype(np.float64) - p1, want, rtol=2e-3, atol=2e-7)
    p, m, v = p1, m1, v1
    for step in range(2, 12):
        pn, m, v = mo.adam_step(p, g, m, v, step, lr)
        assert np.allclose(p.astype(np.float64) - pn, want, rtol=2e-3, atol=2e-7), step
        p = pn
    # the moments are the uncorrected exponential averages
    assert np.allclose(m, (1 - 0.9 ** 11) * g, rtol=1e-5)
    assert np.allclose(v, (1 - 0.999 ** 11) * g.astype(np.float64) ** 2, rtol=1e-4)
