"""Every launch mapping (1, 2 or 4 cooperating wavefronts per 64-scenario
block) gives bit-identical rewards, equal to the oracle; the autotuner picks
one of them."""
import os

import numpy as np
import pytest

from gpu_util import assert_bitwise
from sigp.instance import load_instance, synthetic_uniform
from sigp.rules import EXTRA, NS_TERMINALS, pack_trees, term_mask_from
from sigp.scenarios import gen_durations
from sigp.workloads import random_rules

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


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
    for w in (1, 2, 4):
        print("mapping", name, "waves_per_item", w, flush=True)
        ctx.configure(w)
        assert_bitwise(ctx.evaluate(pk, S, 100, mask), want, "waves_per_item=%d" % w)


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
    w = b.autotune(st.data_ptr(), 128)
    assert w in (1, 2, 4)
    b.run(st.data_ptr(), 128, out.data_ptr())
    torch.cuda.synchronize()
    assert_bitwise(out.cpu().numpy(), want, "tuned batch")
