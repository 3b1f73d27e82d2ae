"""CPU ORACLE -- test infrastructure only: the surrogate network of
sigp_mlp.hip restated in float64 numpy.

forward_backward() follows the device's computation step by step.  With
emulate=True every value the device rounds to BF16 is rounded the same way:
the inputs, the weight matrices W1 and W2, the ReLU outputs and the
back-propagated deltas.  Products and sums are then exact in float64, so the
device's results differ only by its FP32 accumulation.  With emulate=False
it is the plain float64 network, whose gradients are checked against
central finite differences.

adam_step() is the device's optimiser step (adam_kernel: beta1 0.9, beta2
0.999, eps 1e-8, bias-corrected, Kingma & Ba's published update) in float32,
in the kernel's operation order, with the host's powf for the bias
corrections.

Pins: finite differences (tests/test_mlp_oracle.py) for the gradient
formulas; exact integer cases for the BF16 emulation; for adam_step, what
the published update fixes: with zero moments the bias-corrected first step
is lr * g / (|g| + eps), and under a constant gradient every later step is
the same, which a wrong decay rate or a dropped correction breaks.
"""
import ctypes
import ctypes.util
import numpy as np

IN, H = 512, 256
O_W1 = 0
O_B1 = O_W1 + IN * H
O_W2 = O_B1 + H
O_B2 = O_W2 + H * H
O_W3 = O_B2 + H
O_B3 = O_W3 + H
NP = O_B3 + 1


def bf16(a):
    """Round float32 values to bfloat16, nearest even; returns float64."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[O_W1:O_B1].reshape(IN, H), p[O_B1:O_W2], p[O_W2:O_B2].reshape(H, H), p[O_B2:O_W3],
            p[O_W3:O_B3], p[O_B3])


def forward(p, x, emulate=True):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    xq = r(x)
    h1 = xq @ r(W1) + b1
    a1 = r(np.maximum(h1, 0.0))
    h2 = a1 @ r(W2) + b2
    a2 = r(np.maximum(h2, 0.0))
    y = a2 @ w3 + b3
    return y, (xq, h1, a1, h2, a2)


def forward_backward(p, x, t, emulate=True):
    """Returns (mean squared error, gradient vector of NP entries)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    t = np.asarray(t, dtype=np.float64)
    nb = len(t)
    y, (xq, h1, a1, h2, a2) = forward(p, x, emulate)
    loss = float(np.mean((y - t) ** 2))
    dy = 2.0 * (y - t) / nb
    if emulate:   # the device forms dy in float32
        dy = dy.astype(np.float32).astype(np.float64)
    d2 = r(np.where(h2 > 0.0, dy[:, None] * w3[None, :], 0.0))
    d1 = r(np.where(h1 > 0.0, d2 @ r(W2).T, 0.0))
    g = np.zeros(NP)
    g[O_W1:O_B1] = (xq.T @ d1).ravel()
    g[O_B1:O_W2] = d1.sum(axis=0)
    g[O_W2:O_B2] = (a1.T @ d2).ravel()
    g[O_B2:O_W3] = d2.sum(axis=0)
    g[O_W3:O_B3] = a2.T @ dy
    g[O_B3] = dy.sum()
    return loss, g


def loss_only(p, x, t, emulate=False):
    y, _ = forward(p, x, emulate)
    return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))


_libm = ctypes.CDLL(ctypes.util.find_library("m"))
_libm.powf.argtypes = [ctypes.c_float, ctypes.c_float]
_libm.powf.restype = ctypes.c_float


def adam_step(p, g, m, v, step, lr):
    """One Adam step on float32 arrays as adam_kernel computes it; `step` is
    the 1-based step count.  Returns (p, m, v) as float32."""
    f = np.float32
    p, g, m, v = (np.asarray(a, dtype=np.float32) for a in (p, g, m, v))
    c1 = f(1.0) - f(_libm.powf(0.9, float(step)))
    c2 = f(1.0) - f(_libm.powf(0.999, float(step)))
    mk = f(0.9) * m + f(0.1) * g
    vk = f(0.999) * v + f(0.001) * g * g
    pk = p - f(lr) * (mk / c1) / (np.sqrt(vk / c2) + f(1e-8))
    return pk.astype(np.float32), mk.astype(np.float32), vk.astype(np.float32)
