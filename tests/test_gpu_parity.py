"""GPU evaluator (through the C ABI) against the CPU oracle on the same
seeded inputs, bit for bit: random rule sets under every mask family, the
three size buckets, grouped / per-mask index maths and edge cases."""
import os

import numpy as np
import pytest

from gpu_util import assert_bitwise
from sigp.instance import load_instance, synthetic_clustered, synthetic_mixed, synthetic_uniform
from sigp.rules import (EXTRA, FUNCS_RMOD, NS_TERMINALS, SLOT_NAMES, Node, pack_trees,
                        parse_prefix, term_mask_from, to_prefix)
from sigp.scenarios import gen_durations
from sigp.workloads import random_rules

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def _check(oracle, inst, trees, S, mask, label, masks=None, group_size=None, mask_group=None, T=None):
    from sigp.api import FactoryContext
    print("parity", label, "n=%d m=%d P=%d" % (inst.n, inst.m, len(trees)), flush=True)
    ctx = FactoryContext(inst)
    pk = pack_trees(trees)
    T = S.shape[0] if T is None else T
    if masks is None:
        got = ctx.evaluate(pk, S, T, mask)
        want = oracle.evaluate(inst, pk, S, T, mask, ctx.zp, threads=THREADS)
    else:
        got = ctx.evaluate_grouped_masks(pk, S, T, group_size, mask_group, masks)
        want = oracle.evaluate(inst, pk, S, T, masks, ctx.zp, group_size=group_size,
                               mask_group=mask_group, threads=THREADS)
    assert_bitwise(got, want, label)
    return got


def test_extra18_uniform_gamma(oracle):
    inst = synthetic_uniform(n=100, m=2, seed=1)
    trees = random_rules(48, EXTRA, seed=3)
    S = gen_durations(inst.durations, 70, 1_000_000, mode="gamma", shape=4.0)
    _check(oracle, inst, trees, S, term_mask_from(EXTRA), "extra18-uniform")


@pytest.mark.parametrize("name", ["r201", "pr01", "rc105"])
def test_ext29_benchmark_instances(oracle, name):
    inst = load_instance(name)
    trees = random_rules(24, NS_TERMINALS + EXTRA, seed=11, min_depth=2, max_depth=6)
    S = gen_durations(inst.durations, 40, 2_000_000)
    _check(oracle, inst, trees, S, term_mask_from(NS_TERMINALS + EXTRA), "ext29-" + name)


def test_full61_every_slot(oracle):
    """Random trees over all 61 terminals, all slots live (incl. ZTC..ZD2D,
    SKIPLOSS, DETOUR, NNCHAIN3, NSA05, RAND)."""
    inst = load_instance("c105")
    trees = random_rules(40, SLOT_NAMES, seed=5, min_depth=3, max_depth=5)
    S = gen_durations(inst.durations, 20, 77)
    _check(oracle, inst, trees, S, np.ones(61, dtype=np.int64), "full61-c105")


def test_masked_off_slots_read_zero_and_z_partners(oracle):
    """Slots used by a tree but masked off read 0.0; REGRET/REACH2/ATRISK are
    live when only their Z partner is in the mask."""
    inst = load_instance("rc101")
    mask = term_mask_from(["SCORE", "ZWLOSS", "ZREACH", "ZATRISK", "NS"])
    trees = [parse_prefix(t) for t in [
        "(+ REGRET (* 0.5 REACH2))", "(- ATRISK ZATRISK)", "(max ZWLOSS (min NSROB ZREACH))",
        "(+ (* HARVEST 2.0) (- SCORE BOLSAO))", "(/ NS (+ REGRET 1.0))"]]
    S = gen_durations(inst.durations, 30, 5)
    _check(oracle, inst, trees, S, mask, "mask-semantics")


def test_rmod_rules(oracle):
    """256 rules that each hold at least one rmod node x 256 scenarios: every
    simulation calls pow at every decision, so a last-bit difference from
    glibc's pow that flipped any argmax would show.  The oracle computes
    only the slots each rule reads (equal to the full evaluation bit for
    bit, tests/test_oracle_checks.py::test_pruned_equals_full)."""
    from oracle import need_rows_for
    from sigp.api import FactoryContext
    inst = load_instance("c101")
    pool = random_rules(768, NS_TERMINALS + EXTRA, seed=9, min_depth=2, max_depth=5,
                        functions=FUNCS_RMOD)
    trees = [t for t in pool if "rmod" in to_prefix(t)][:256]
    assert len(trees) == 256
    S = gen_durations(inst.durations, 256, 3)
    mask = term_mask_from(NS_TERMINALS + EXTRA)
    pk = pack_trees(trees)
    print("parity rmod n=%d P=%d T=%d" % (inst.n, pk.n_rules, S.shape[0]), flush=True)
    ctx = FactoryContext(inst)
    got = ctx.evaluate(pk, S, S.shape[0], mask)
    want = oracle.evaluate(inst, pk, S, S.shape[0], mask, ctx.zp, need_rows=need_rows_for(pk, mask),
                           threads=THREADS)
    assert_bitwise(got, want, "rmod")


def test_large_bucket_and_clustered(oracle):
    inst = synthetic_mixed(n=288, m=4, seed=2)
    trees = random_rules(8, EXTRA, seed=1)
    S = gen_durations(inst.durations, 64, 10, mode="gamma", shape=2.0)
    _check(oracle, inst, trees, S, term_mask_from(EXTRA), "n288")
    inst2 = synthetic_clustered(n=100, m=3, seed=4)
    S2 = gen_durations(inst2.durations, 33, 10)
    _check(oracle, inst2, random_rules(16, NS_TERMINALS + EXTRA, seed=2), S2,
           term_mask_from(NS_TERMINALS + EXTRA), "clustered")


def test_largest_bucket(oracle):
    inst = synthetic_uniform(n=400, m=3, seed=6, width=(20.0, 200.0))
    trees = random_rules(4, EXTRA, seed=4, max_depth=3)
    S = gen_durations(inst.durations, 20, 10, mode="gamma")
    _check(oracle, inst, trees, S, term_mask_from(EXTRA), "n400")


def test_grouped_masks_index_maths(oracle):
    """Tree p reads scenario group p // group_size and mask row p // mask_group
    (engine/factory.cuh:98-102)."""
    inst = load_instance("r105")
    trees = random_rules(24, NS_TERMINALS + EXTRA, seed=21)
    G, T = 3, 9
    S = np.concatenate([gen_durations(inst.durations, T, 100 * g) for g in range(G)])
    masks = np.stack([term_mask_from(NS_TERMINALS), term_mask_from(NS_TERMINALS + EXTRA[:9]),
                      term_mask_from(NS_TERMINALS + EXTRA)])
    _check(oracle, inst, trees, S, None, "grouped-masks", masks=masks, group_size=8,
           mask_group=8, T=T)


def test_partial_waves_and_deep_stack(oracle):
    inst = load_instance("pr02")
    deep = Node("REGRET")
    for k in range(20):   # right-leaning chain: stack depth 21 (32-deep variant)
        deep = Node("+", (Node(["SCORE", "NSTW", "TA", "REACH1"][k % 4]), deep))
    trees = [deep, parse_prefix("SCORE"), parse_prefix("(- 0.000 TA)")]
    for T in (1, 63, 65, 130):
        S = gen_durations(inst.durations, T, 1000 + T)
        _check(oracle, inst, trees, S, term_mask_from(NS_TERMINALS + EXTRA), "T=%d" % T)


def test_tiny_and_degenerate_instances(oracle):
    from sigp.api import FactoryContext
    one = synthetic_uniform(n=1, m=2, seed=3)
    S = gen_durations(one.durations, 5, 1)
    _check(oracle, one, [parse_prefix("(+ SCORE NS)")], S, term_mask_from(NS_TERMINALS), "n=1")
    none = synthetic_uniform(n=5, m=0, seed=3)
    ctx = FactoryContext(none)
    pk = pack_trees([parse_prefix("SCORE")])
    out = ctx.evaluate(pk, gen_durations(none.durations, 3, 1), 3, term_mask_from(NS_TERMINALS))
    assert np.array_equal(out, np.zeros(3))


def test_means_and_batch_api():
    from sigp.api import FactoryContext
    inst = load_instance("c201")
    trees = random_rules(10, EXTRA, seed=2)
    pk = pack_trees(trees)
    S = gen_durations(inst.durations, 37, 9)
    ctx = FactoryContext(inst)
    pairs, means = ctx.evaluate_means(pk, S, 37, term_mask_from(EXTRA))
    want = np.array([sum(pairs[p * 37:(p + 1) * 37]) / 37 for p in range(10)])
    assert_bitwise(means, want, "means")


def test_invalid_arguments_rejected():
    from sigp._lib import SigpError
    from sigp.api import FactoryContext
    from sigp.rules import Packed
    inst = load_instance("pr01")
    ctx = FactoryContext(inst)
    S = gen_durations(inst.durations, 2, 1)
    mask = term_mask_from(NS_TERMINALS)
    i32 = lambda v: np.array(v, dtype=np.int32)
    bad = [
        Packed(i32([-1, 0]), i32([0, 0, 0, 0]), i32([0]), i32([2])),          # underflow
        Packed(i32([-1, -1, 9]), i32([0, 0, 1, 0, 0, 1]), i32([0]), i32([3])),  # opcode 9
        Packed(i32([-1]), i32([61, 0]), i32([0]), i32([1])),                  # slot 61
        Packed(i32([-1, -1]), i32([0, 0, 1, 0]), i32([0]), i32([2])),          # two values left
    ]
    for pk in bad:
        with pytest.raises(SigpError):
            ctx.evaluate(pk, S, 2, mask)
    with pytest.raises(SigpError):   # fewer scenario rows than T
        ctx.evaluate(pack_trees([parse_prefix("SCORE")]), S, 3, mask)
