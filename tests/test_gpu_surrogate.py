"""The surrogate on the device: the BF16 matrix-core network against its
CPU oracle (exactly on integer cases, to FP32 accumulation otherwise), its
training, and the phenotype kernel against the CPU oracle's priorities."""
import numpy as np
import pytest

import mlp_oracle as mo
from sigp.rules import EXTRA, NS_TERMINALS, pack_trees, term_mask_from
from sigp.workloads import plain_small, random_rules

pytestmark = pytest.mark.gpu


def test_forward_exact_on_integers():
    """Small integers are exact in BF16 and their sums exact in FP32, so the
    device output must equal the oracle's bit for bit: pins the MFMA operand
    layout and every GEMM index."""
    from sigp.surrogate import Mlp
    rng = np.random.default_rng(11)

    def sparse(n, keep):
        return rng.integers(-1, 2, n) * (rng.random(n) < keep)

    p = np.zeros(mo.NP)
    p[mo.O_W1:mo.O_B1] = rng.integers(-1, 2, mo.IN * mo.H)
    p[mo.O_B1:mo.O_W2] = rng.integers(-3, 4, mo.H)
    p[mo.O_W2:mo.O_B2] = sparse(mo.H * mo.H, 0.3)
    p[mo.O_B2:mo.O_W3] = rng.integers(-3, 4, mo.H)
    p[mo.O_W3:mo.O_B3] = sparse(mo.H, 0.3)
    p[mo.O_B3] = 5
    x = rng.integers(0, 2, (70, mo.IN)).astype(np.float32)   # 70 rows: a padded batch
    want, (xq, h1, a1, h2, a2) = mo.forward(p, x, emulate=True)
    # every partial sum is an integer below 2^24, so FP32 accumulation is exact
    w3 = p[mo.O_W3:mo.O_B3]
    bound = max(np.abs(h1).max(), np.abs(h2).max(), (np.abs(a2) @ np.abs(w3)).max() + 5)
    assert bound < 2 ** 24, bound
    net = Mlp(seed=1)
    net.set_params(p)
    got = net.predict(x)
    print("largest partial sum", bound, flush=True)
    assert np.array_equal(got.astype(np.float64), want)


def test_gradients_match_oracle():
    from sigp.surrogate import Mlp
    rng = np.random.default_rng(5)
    net = Mlp(seed=3)
    p = net.params().astype(np.float64)
    p[mo.O_B1:mo.O_W2] = rng.normal(0, 0.05, mo.H)
    p[mo.O_B2:mo.O_W3] = rng.normal(0, 0.05, mo.H)
    net.set_params(p)
    p = net.params().astype(np.float64)
    x = (rng.integers(0, 8, (100, mo.IN)) / 7.0).astype(np.float32)
    t = rng.normal(size=100).astype(np.float32)
    g, loss = net.gradients(x, t)
    want_loss, want = mo.forward_backward(p, x, t, emulate=True)
    assert abs(loss - want_loss) <= 1e-4 * max(1.0, want_loss)
    for lo, hi, name in [(mo.O_W1, mo.O_B1, "W1"), (mo.O_B1, mo.O_W2, "b1"), (mo.O_W2, mo.O_B2, "W2"),
                         (mo.O_B2, mo.O_W3, "b2"), (mo.O_W3, mo.O_B3, "w3"), (mo.O_B3, mo.NP, "b3")]:
        a, b = g[lo:hi].astype(np.float64), want[lo:hi]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(name, "relative error %.2e" % err, flush=True)
        assert err < 2e-3, name


def test_adam_steps_match_oracle():
    """Two training steps on one row (batch 1, so each epoch is one step and
    the row order is fixed) equal the oracle's Adam (float32, kernel order)
    applied to the device's own gradients, bit for bit: pins adam_kernel,
    its bias corrections and the refresh of the BF16 weight copies the
    second gradient is taken with."""
    from sigp.surrogate import Mlp
    rng = np.random.default_rng(21)
    x = (rng.integers(0, 8, (1, mo.IN)) / 7.0).astype(np.float32)
    t = np.array([0.7], dtype=np.float32)
    lr = np.float32(1e-3)
    net = Mlp(seed=5)
    p0 = net.params()
    g0, _ = net.gradients(x, t)
    net.train(x, t, epochs=2, batch=1, lr=float(lr), seed=0)
    got = net.params()
    z = np.zeros_like(p0)
    p1, m1, v1 = mo.adam_step(p0, g0, z, z, 1, lr)
    probe = Mlp(seed=6)
    probe.set_params(p1)
    g1, _ = probe.gradients(x, t)
    p2, _, _ = mo.adam_step(p1, g1, m1, v1, 2, lr)
    moved = np.count_nonzero(p2 != p0)
    print("parameters moved by two steps: %d of %d" % (moved, p0.size), flush=True)
    assert moved > p0.size // 4
    assert np.array_equal(got.view(np.uint32), p2.view(np.uint32))


def test_training_fits_a_known_function():
    from sigp.surrogate import Mlp, spearman
    rng = np.random.default_rng(9)
    x = (rng.integers(0, 8, (4096, mo.IN)) / 7.0).astype(np.float32)
    w = rng.normal(size=mo.IN) / np.sqrt(mo.IN)
    t = (np.tanh(x @ w) + 0.3 * np.sin(3 * x[:, 0])).astype(np.float32)
    t = (t - t.mean()) / t.std()
    net = Mlp(seed=2)
    losses = net.train(x[:3584], t[:3584], epochs=25, batch=256, lr=1e-3, seed=4)
    print("epoch losses", np.round(losses, 4), flush=True)
    assert losses[-1] < 0.25 * losses[0]
    rho = spearman(net.predict(x[3584:]), t[3584:])
    print("held-out spearman %.3f" % rho, flush=True)
    assert rho > 0.8


def test_phenotype_matches_oracle(oracle):
    from sigp.api import FactoryContext
    from sigp.surrogate import make_situations, phenotype
    wl = plain_small(n_rules=4, n_scen=2)
    inst = wl.inst
    trees = random_rules(12, NS_TERMINALS + EXTRA, seed=19, min_depth=1, max_depth=5)
    pk = pack_trees(trees)
    mask = term_mask_from(NS_TERMINALS + EXTRA)
    sit = make_situations(inst, seed=5)
    ctx = FactoryContext(inst)
    got = phenotype(ctx, pk, mask, sit)
    for p in range(pk.n_rules):
        o, L = int(pk.offsets[p]), int(pk.lengths[p])
        ops, args = pk.ops[o:o + L], pk.args[2 * o:2 * (o + L)]
        for s in range(sit.n_sit):
            _, pri = oracle.pick_best(inst, int(sit.cur[s]), float(sit.ct[s]), int(sit.remaining[s]), ops,
                                      args, sit.visited[s], mask, ctx.zp)
            pc = np.maximum(np.nan_to_num(pri[sit.cand[s]], nan=-1e300), -1e300)
            want = [sum(1 for k in range(len(pc)) if pc[k] > pc[j] or (pc[k] == pc[j] and k < j))
                    for j in range(len(pc))]
            assert list(got[p, s]) == want, (p, s, list(got[p, s]), want)
