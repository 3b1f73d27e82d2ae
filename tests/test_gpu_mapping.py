"""Every launch mapping gives bit-identical rewards, equal to the oracle:
scenario lanes with 1, 2 or 4 cooperating wavefronts per 64-scenario block,
candidate lanes with 1, 2 or 4 wavefronts per (rule, scenario) simulation,
and segmented candidate lanes (2 simulations per wavefront); the
autotuner picks one of them."""
import os

import numpy as np
import pytest

from gpu_util import assert_bitwise
from sigp.instance import load_instance, synthetic_mixed, synthetic_uniform
from sigp.rules import EXTRA, NS_TERMINALS, pack_trees, term_mask_from
from sigp.scenarios import gen_durations
from sigp.workloads import random_rules

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
MAPS = [(1, w) for w in (1, 2, 4)] + [(2, w) for w in (1, 2, 4)] + [(3, 1)]   # (mapping, waves)


@pytest.mark.parametrize("name", ["c102", "syn"])
def test_mappings_agree_with_oracle(oracle, name):
    from sigp.api import FactoryContext
    inst = synthetic_uniform(n=100, m=2, seed=9) if name == "syn" else load_instance(name)
    trees = random_rules(40, NS_TERMINALS + EXTRA, seed=13, min_depth=1, max_depth=5)
    pk = pack_trees(trees)
    mask = term_mask_from(NS_TERMINALS + EXTRA)
    S = gen_durations(inst.durations, 100, 5, mode="gamma")
    ctx = FactoryContext(inst)
    want = oracle.evaluate(inst, pk, S, 100, mask, ctx.zp, threads=THREADS)
    for m, w in MAPS:
        print("mapping", name, m, "waves", w, flush=True)
        ctx.configure(w, m)
        assert_bitwise(ctx.evaluate(pk, S, 100, mask), want, "mapping %d waves %d" % (m, w))


def test_gp_loop_shape_one_scenario_per_group(oracle):
    """The reference GP loop's launch: G seeds x population rules, one
    scenario each, rule p reading group p / population (engine/gp_engine.cu:
    155-164)."""
    from sigp.api import FactoryContext
    inst = load_instance("r101")
    G, pop = 4, 48
    trees = random_rules(G * pop, NS_TERMINALS + EXTRA, seed=21, min_depth=1, max_depth=6)
    pk = pack_trees(trees)
    mask = term_mask_from(NS_TERMINALS + EXTRA)
    S = np.concatenate([gen_durations(inst.durations, 1, g * 10000 + 7) for g in range(G)])
    ctx = FactoryContext(inst)
    want = oracle.evaluate(inst, pk, S, 1, mask, ctx.zp, group_size=pop, threads=THREADS)
    for m, w in [(0, 0)] + MAPS:
        print("gp-loop shape mapping", m, "waves", w, flush=True)
        ctx.configure(w, m)
        got = ctx.evaluate_grouped(pk, S, G, 1, pop, mask)
        assert_bitwise(got, want, "T=1 mapping %d waves %d" % (m, w))


def test_candidate_lanes_more_candidates_than_threads(oracle):
    """n = 288 with 256 threads: threads take two candidates each; in 32-lane
    segments, up to 9 each (T = 3: the second wavefront's second segment
    idles)."""
    from sigp.api import FactoryContext
    inst = synthetic_mixed(n=288, m=4, seed=5)
    trees = random_rules(12, EXTRA, seed=8, min_depth=1, max_depth=4)
    pk = pack_trees(trees)
    mask = term_mask_from(EXTRA)
    S = gen_durations(inst.durations, 3, 11, mode="gamma", shape=2.0)
    ctx = FactoryContext(inst)
    want = oracle.evaluate(inst, pk, S, 3, mask, ctx.zp, threads=THREADS)
    for m, w in [(2, 1), (2, 2), (2, 4), (3, 1)]:
        print("n=288 candidate lanes mapping", m, "waves", w, flush=True)
        ctx.configure(w, m)
        assert_bitwise(ctx.evaluate(pk, S, 3, mask), want, "n=288 mapping %d waves %d" % (m, w))


def test_autotune_and_device_batch():
    import torch
    from sigp.api import FactoryContext
    inst = load_instance("r102")
    trees = random_rules(64, EXTRA, seed=3)
    pk = pack_trees(trees)
    mask = term_mask_from(EXTRA)
    S = gen_durations(inst.durations, 128, 77)
    ctx = FactoryContext(inst)
    want = ctx.evaluate(pk, S, 128, mask)
    b = ctx.batch(pk, 128, mask)
    dev = torch.device("cuda:0")
    st = torch.from_numpy(S).to(dev)
    out = torch.zeros(64 * 128, dtype=torch.float64, device=dev)
    w, m = b.autotune(st.data_ptr(), 128)
    assert w in (1, 2, 4) and m in (1, 2, 3)
    b.run(st.data_ptr(), 128, out.data_ptr())
    torch.cuda.synchronize()
    assert_bitwise(out.cpu().numpy(), want, "tuned batch")


@pytest.mark.parametrize("name", ["c105", "mixed288"])
def test_every_kernel_variant_under_every_mapping(oracle, name):
    """All 61 slots with rmod (the RMOD and RARE variants), n = 100 and
    n = 288 (both segmented builds), an odd T, then a one-customer and a
    no-vehicle instance: every mapping equals the oracle."""
    from sigp.api import FactoryContext
    from sigp.rules import FUNCS_RMOD, SLOT_NAMES, parse_prefix
    inst = load_instance(name) if name == "c105" else synthetic_mixed(n=288, m=4, seed=2)
    trees = random_rules(16, SLOT_NAMES, seed=17, min_depth=2, max_depth=4, functions=FUNCS_RMOD)
    pk = pack_trees(trees)
    mask = np.ones(61, dtype=np.int64)
    S = gen_durations(inst.durations, 5, 303, mode="gamma")
    ctx = FactoryContext(inst)
    want = oracle.evaluate(inst, pk, S, 5, mask, ctx.zp, threads=THREADS)
    one = synthetic_uniform(n=1, m=2, seed=3)
    S1 = gen_durations(one.durations, 3, 1)
    c1 = FactoryContext(one)
    p1 = pack_trees([parse_prefix("(+ SCORE NS)")])
    want1 = oracle.evaluate(one, p1, S1, 3, term_mask_from(NS_TERMINALS), c1.zp, threads=THREADS)
    none = synthetic_uniform(n=5, m=0, seed=3)
    c0 = FactoryContext(none)
    for m, w in [(1, 1), (2, 1), (3, 1)]:
        print("variants", name, "mapping", m, flush=True)
        for c in (ctx, c1, c0):
            c.configure(w, m)
        assert_bitwise(ctx.evaluate(pk, S, 5, mask), want, "%s mapping %d" % (name, m))
        assert_bitwise(c1.evaluate(p1, S1, 3, term_mask_from(NS_TERMINALS)), want1, "n=1 mapping %d" % m)
        out = c0.evaluate(pack_trees([parse_prefix("SCORE")]), gen_durations(none.durations, 3, 1), 3,
                          term_mask_from(NS_TERMINALS))
        assert np.array_equal(out, np.zeros(3)), "m=0 mapping %d" % m


def test_scenario_order_changes_nothing(oracle):
    """Scenario lanes with and without the scenario order
    (sigp_batch_order_scenarios): identical, and equal to the oracle, on one
    instance (T = 130: a partial block), on grouped rows (3 groups) and on a
    multi-instance batch."""
    import torch
    from sigp.api import FactoryContext, batch_multi
    dev = torch.device("cuda:0")
    inst = load_instance("r102")
    trees = random_rules(24, NS_TERMINALS + EXTRA, seed=29, min_depth=2, max_depth=5)
    pk = pack_trees(trees)
    mask = term_mask_from(NS_TERMINALS + EXTRA)
    ctx = FactoryContext(inst)
    ctx.configure(4, 1)

    def run(b, S, n_out, n_rows=None):
        st = torch.from_numpy(np.ascontiguousarray(S)).to(dev)
        out = torch.zeros(n_out, dtype=torch.float64, device=dev)
        b.run(st.data_ptr(), S.shape[0] if n_rows is None else n_rows, out.data_ptr())
        torch.cuda.synchronize()
        return out.cpu().numpy()

    T = 130
    S = gen_durations(inst.durations, T, 4242, mode="gamma")
    want = oracle.evaluate(inst, pk, S, T, mask, ctx.zp, threads=THREADS)
    b = ctx.batch(pk, T, mask)
    assert b.order_scenarios(True)
    on = run(b, S, 24 * T)
    assert not b.order_scenarios(False)
    off = run(b, S, 24 * T)
    assert_bitwise(on, want, "ordered")
    assert_bitwise(off, want, "index order")

    G, gs, T = 3, 8, 70
    S = np.concatenate([gen_durations(inst.durations, T, 500 + g) for g in range(G)])
    want = oracle.evaluate(inst, pk, S, T, mask, ctx.zp, group_size=gs, threads=THREADS)
    assert_bitwise(ctx.evaluate_grouped(pk, S, G, T, gs, mask), want, "grouped, ordered")

    insts = [load_instance("c101"), synthetic_mixed(n=150, m=3, seed=4), synthetic_uniform(n=60, m=2, seed=6)]
    ctxs = [FactoryContext(i) for i in insts]
    T = 64
    blocks = [gen_durations(i.durations, T, 77 + k, mode="gamma") for k, i in enumerate(insts)]
    want = np.concatenate([oracle.evaluate(i, pk, B, T, mask, c.zp, threads=THREADS)
                           for i, c, B in zip(insts, ctxs, blocks)])
    bm = batch_multi(ctxs, pk, T, mask)
    bm.configure(4, 1)
    assert bm.order_scenarios(True)
    flat = np.concatenate([B.ravel() for B in blocks])   # instance blocks back to back
    assert_bitwise(run(bm, flat, len(insts) * 24 * T, len(insts) * T), want, "multi, ordered")
    ctx.configure(0, 0)


def test_skip_repeats_changes_nothing(oracle):
    """sigp_instance_skip_repeats: rule sets with repeats (within a scenario
    group and across groups, under two mask rows, on one and on several
    instances) give the oracle's rewards bit for bit, skipped or not."""
    from sigp.api import FactoryContext, evaluate_multi
    inst = load_instance("rc102")
    base = random_rules(12, NS_TERMINALS + EXTRA, seed=41, min_depth=1, max_depth=4)
    trees = base + base[:5] + base[2:4] + base[:1]      # 20 rules, 8 repeats
    pk = pack_trees(trees)
    mask = term_mask_from(NS_TERMINALS + EXTRA)
    masks = np.stack([mask, term_mask_from(NS_TERMINALS)])
    ctx = FactoryContext(inst)
    T = 70
    S = gen_durations(inst.durations, T, 9090, mode="gamma")
    want = oracle.evaluate(inst, pk, S, T, mask, ctx.zp, threads=THREADS)
    G, gs = 2, 10                                        # groups of 10 rules: repeats on both sides of the cut
    SG = np.concatenate([gen_durations(inst.durations, T, 100 + g) for g in range(G)])
    want_g = oracle.evaluate(inst, pk, SG, T, masks, ctx.zp, group_size=gs, mask_group=10, threads=THREADS)
    insts = [inst, synthetic_mixed(n=150, m=3, seed=8)]
    ctxs = [ctx, FactoryContext(insts[1])]
    blocks = [gen_durations(i.durations, 24, 55 + k) for k, i in enumerate(insts)]
    want_m = np.concatenate([oracle.evaluate(i, pk, B, 24, mask, c.zp, threads=THREADS)
                             for i, c, B in zip(insts, ctxs, blocks)])
    for skip in (True, False):
        for c in ctxs:
            c.skip_repeats(skip)
        for m, w in [(0, 0), (1, 4), (2, 1)]:
            ctx.configure(w, m)
            print("skip repeats", skip, "mapping", m, flush=True)
            assert_bitwise(ctx.evaluate(pk, S, T, mask), want, "skip %s mapping %d" % (skip, m))
            assert_bitwise(ctx.evaluate_grouped_masks(pk, SG, T, gs, 10, masks), want_g,
                           "grouped, skip %s mapping %d" % (skip, m))
        ctx.configure(0, 0)
        assert_bitwise(evaluate_multi(ctxs, pk, blocks, 24, mask), want_m, "multi, skip %s" % skip)
        means = ctx.evaluate_means(pk, S, T, mask)[1]
        assert np.array_equal(means, want.reshape(len(trees), T).mean(axis=1)) or \
            np.all(means == np.array([want[p * T:(p + 1) * T].sum() / T for p in range(len(trees))]))
    ctx.skip_repeats(False)
