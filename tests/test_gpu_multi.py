"""Multi-instance batches (sigp_batch_create_multi / sigp_evaluate_multi):
every rule on every instance, instances of uneven size in one launch per
size class.  Each instance's block of results must equal the oracle's
single-instance evaluation bit for bit, under every mapping and after
autotuning."""
import os

import numpy as np
import pytest

from gpu_util import assert_bitwise
from sigp.instance import load_instance, synthetic_clustered, synthetic_mixed, synthetic_uniform
from sigp.rules import EXTRA, NS_TERMINALS, pack_trees, term_mask_from
from sigp.scenarios import gen_durations
from sigp.workloads import random_rules

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def _instances():
    return [synthetic_uniform(n=48, m=1, seed=3, width=(10.0, 200.0)),
            synthetic_clustered(n=144, m=2, seed=4, width=(10.0, 200.0)),
            load_instance("c101"),
            synthetic_mixed(n=288, m=4, seed=5),
            synthetic_uniform(n=96, m=3, seed=6, width=(10.0, 200.0))]


def _expected(oracle, ctxs, insts, pk, blocks, T, mask):
    return np.concatenate([oracle.evaluate(inst, pk, B, T, mask, ctx.zp, threads=THREADS)
                           for inst, ctx, B in zip(insts, ctxs, blocks)])


def test_multi_instance_matches_oracle_every_mapping(oracle):
    from sigp.api import FactoryContext, evaluate_multi
    insts = _instances()
    ctxs = [FactoryContext(i) for i in insts]
    trees = random_rules(20, NS_TERMINALS + EXTRA, seed=31, min_depth=1, max_depth=5)
    pk = pack_trees(trees)
    mask = term_mask_from(NS_TERMINALS + EXTRA)
    T = 24
    blocks = [gen_durations(i.durations, T, 1000 + k, mode="gamma") for k, i in enumerate(insts)]
    want = _expected(oracle, ctxs, insts, pk, blocks, T, mask)
    for m, w in [(0, 0), (1, 1), (1, 4), (2, 1), (2, 4), (3, 1)]:
        print("multi mapping", m, "waves", w, flush=True)
        ctxs[0].configure(w, m)          # the batch takes its first instance's default
        got = evaluate_multi(ctxs, pk, blocks, T, mask)
        assert_bitwise(got, want, "multi mapping %d waves %d" % (m, w))
    ctxs[0].configure(0, 0)


def test_multi_instance_device_batch_autotuned_with_means(oracle):
    import torch
    from sigp.api import FactoryContext, batch_multi
    insts = _instances()[:4]
    ctxs = [FactoryContext(i) for i in insts]
    trees = random_rules(16, EXTRA, seed=5, min_depth=1, max_depth=6)
    pk = pack_trees(trees)
    mask = term_mask_from(EXTRA)
    T = 70
    blocks = [gen_durations(i.durations, T, 77 + k, mode="gamma") for k, i in enumerate(insts)]
    want = _expected(oracle, ctxs, insts, pk, blocks, T, mask)
    b = batch_multi(ctxs, pk, T, mask)
    dev = torch.device("cuda:0")
    flat = torch.from_numpy(np.concatenate([B.ravel() for B in blocks])).to(dev)
    K, P = len(insts), pk.n_rules
    out = torch.zeros(K * P * T, dtype=torch.float64, device=dev)
    means = torch.zeros(K * P, dtype=torch.float64, device=dev)
    w, m = b.autotune(flat.data_ptr(), K * T)
    print("autotuned", w, m, flush=True)
    b.run(flat.data_ptr(), K * T, out.data_ptr(), means.data_ptr())
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert_bitwise(got, want, "tuned multi batch")
    R = want.reshape(K * P, T)
    assert_bitwise(means.cpu().numpy(), np.array([sum(r.tolist()) / T for r in R]), "multi means")
    with pytest.raises(Exception):
        b.run(flat.data_ptr(), K * T - 1, out.data_ptr())   # too few rows: rejected on the host


def test_multi_instance_one_launch(oracle):
    """Every size class in one launch (sigp_instance_one_launch on every
    instance): the same results as the per-class launches and the oracle,
    under scenario, candidate and segmented lanes; with one instance not
    asking, the batch keeps its per-class launches (same results again)."""
    from sigp.api import FactoryContext, evaluate_multi
    insts = _instances()
    ctxs = [FactoryContext(i) for i in insts]
    trees = random_rules(20, NS_TERMINALS + EXTRA, seed=37, min_depth=1, max_depth=5)
    pk = pack_trees(trees)
    mask = term_mask_from(NS_TERMINALS + EXTRA)
    T = 70
    blocks = [gen_durations(i.durations, T, 3000 + k, mode="gamma") for k, i in enumerate(insts)]
    want = _expected(oracle, ctxs, insts, pk, blocks, T, mask)
    for c in ctxs:
        c.one_launch(True)
    for m, w in [(0, 0), (1, 4), (2, 1), (3, 1)]:
        print("one launch, mapping", m, "waves", w, flush=True)
        ctxs[0].configure(w, m)
        assert_bitwise(evaluate_multi(ctxs, pk, blocks, T, mask), want, "one launch mapping %d waves %d" % (m, w))
    ctxs[0].configure(0, 0)
    ctxs[2].one_launch(False)
    assert_bitwise(evaluate_multi(ctxs, pk, blocks, T, mask), want, "per-class launches")
