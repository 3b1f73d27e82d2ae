"""The benchmark workload itself (1 uniform instance n=100 m=2, 1,024
EXTRA-18 rules x 500 Gamma(4) scenarios): a slice checked against the oracle
bit for bit, then the full size through properties that do not depend on
size, plus a random sample of full-size pairs against the oracle."""
import os

import numpy as np
import pytest

from gpu_util import assert_bitwise
from sigp.rules import pack_trees

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def wl():
    from sigp.workloads import plain_small
    return plain_small()


@pytest.fixture(scope="module")
def ctx(wl):
    from sigp.api import FactoryContext
    return FactoryContext(wl.inst)


@pytest.fixture(scope="module")
def full(wl, ctx):
    print("full workload", wl.name, flush=True)
    return ctx.evaluate_means(wl.packed, wl.sampled, wl.n_scen, wl.mask)


def test_slice_bitwise(oracle, wl, ctx):
    trees = wl.trees[::16]                       # 64 rules, every depth/shape class
    pk = pack_trees(trees)
    S = wl.sampled[:96]
    got = ctx.evaluate(pk, S, 96, wl.mask)
    want = oracle.evaluate(wl.inst, pk, S, 96, wl.mask, ctx.zp, threads=THREADS)
    assert_bitwise(got, want, "slice")


def test_full_size_sampled_pairs(oracle, wl, ctx, full):
    pairs, _ = full
    rng = np.random.default_rng(1)
    ps = rng.choice(wl.packed.n_rules, 24, replace=False)
    ss = rng.choice(wl.n_scen, 6, replace=False)
    pk = pack_trees([wl.trees[p] for p in ps])
    S = wl.sampled[ss]
    want = oracle.evaluate(wl.inst, pk, S, len(ss), wl.mask, ctx.zp, threads=THREADS)
    got = pairs.reshape(wl.packed.n_rules, wl.n_scen)[np.ix_(ps, ss)].ravel()
    assert_bitwise(got, want, "sampled pairs")


def test_full_size_properties(wl, ctx, full):
    pairs, means = full
    P, T = wl.packed.n_rules, wl.n_scen
    R = pairs.reshape(P, T)
    # rewards are sums of integer scores, within [0, sum of scores]
    assert np.all(R == np.round(R)) and R.min() >= 0 and R.max() <= wl.inst.scores.sum()
    # per-rule means: ordered sum over scenarios / T
    want = np.array([sum(R[p].tolist()) / T for p in range(P)])
    assert_bitwise(means, want, "means")
    # determinism: a second launch is bit-identical
    again = ctx.evaluate(wl.packed, wl.sampled, T, wl.mask)
    assert_bitwise(again, pairs, "rerun")


def test_scenario_and_rule_permutation(wl, ctx, full):
    """Common random numbers: permuting scenario rows permutes every rule's
    rewards; permuting rules permutes rows of the result."""
    pairs, _ = full
    P, T = wl.packed.n_rules, wl.n_scen
    rng = np.random.default_rng(5)
    sp = rng.permutation(T)
    rp = rng.permutation(P)[:256]
    pk = pack_trees([wl.trees[p] for p in rp])
    got = ctx.evaluate(pk, wl.sampled[sp], T, wl.mask).reshape(len(rp), T)
    want = pairs.reshape(P, T)[rp][:, sp]
    assert_bitwise(got, want, "permutation")
