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
