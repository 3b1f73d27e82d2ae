"""Two benchmark configurations checked at their full size on the GPU.

* The multi-instance batch (40 uneven instances, n 48-288, m 1-4, x 256
  rules of up to 127 nodes x 500 Gamma(4) scenarios, one batch): every
  reward integer-valued and within [0, sum of the instance's scores]; the
  means are the rewards summed in scenario order over T; three instances
  (the smallest, a middle one and the largest) equal single-instance
  batches of the same rules bit for bit; and a random sample of (instance,
  rule, scenario) triples from each size class equals the oracle.
* The GP loop on the device (clustered n = 100, m = 3, population 1,024,
  50 Gamma(4) training scenarios per generation seeded run*10000 +
  generation as engine/gp_engine.cu:69, 50 generations, one HIP graph per
  generation), run in three calls; at generations 17 and 38 the population
  and the generator state are read back and the generation is replayed on
  the host: its fitness by the oracle (bitwise to the device's), then the
  host operators' gp_step (sigp/gp.py, pinned by the reference's gp.txt)
  from the device's generator state, whose offspring must be the device's
  next population tree for tree.  The final best rule's 500-scenario
  held-out mean is the oracle's.
"""
import os
import random

import numpy as np
import pytest

from gpu_util import assert_bitwise
from sigp import gp as G
from sigp.rules import EXTRA, FUNCTIONS, NS_TERMINALS, pack_trees, term_mask_from, to_prefix
from sigp.scenarios import gen_durations

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def _means(pairs, T):
    """Per-rule means summed in scenario order, as the reference's callers
    do (engine/gp_engine.cu:76-81)."""
    return np.cumsum(pairs.reshape(-1, T), axis=1)[:, -1] / T


def test_batched_uneven_full_size(oracle):
    import torch
    from oracle import need_rows_for
    from sigp.api import FactoryContext, batch_multi
    from sigp.workloads import batched_uneven
    wl = batched_uneven()
    K, P, T = len(wl.insts), wl.packed.n_rules, wl.n_scen
    print("batched full size: %d instances x %d rules x %d scenarios" % (K, P, T), flush=True)
    ctxs = [FactoryContext(i) for i in wl.insts]
    batch = batch_multi(ctxs, wl.packed, T, wl.mask)
    dev = torch.device("cuda:0")
    rows = torch.from_numpy(np.concatenate([b.ravel() for b in wl.blocks])).to(dev)
    pairs_d = torch.zeros(K * P * T, dtype=torch.float64, device=dev)
    means_d = torch.zeros(K * P, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    batch.run(rows.data_ptr(), K * T, pairs_d.data_ptr(), means_d.data_ptr(), stream.cuda_stream)
    torch.cuda.synchronize()
    pairs = pairs_d.cpu().numpy().reshape(K, P, T)
    means = means_d.cpu().numpy().reshape(K, P)
    assert np.all(np.isfinite(pairs)) and np.all(pairs == np.round(pairs))
    for k, inst in enumerate(wl.insts):
        assert pairs[k].min() >= 0.0 and pairs[k].max() <= inst.scores.sum(), k
    assert_bitwise(means.ravel(), _means(pairs, T), "means")
    assert pairs.max() > 0
    # three instances against single-instance batches
    ns = [i.n for i in wl.insts]
    picks = sorted({int(np.argmin(ns)), int(np.argsort(ns)[K // 2]), int(np.argmax(ns))})
    for k in picks:
        print("  single-instance batch for instance %d (n=%d m=%d)" % (k, ns[k], wl.insts[k].m), flush=True)
        single = ctxs[k].evaluate(wl.packed, wl.blocks[k], T, wl.mask)
        assert_bitwise(single, pairs[k].ravel(), "instance %d alone" % k)
    # sampled triples of each size class against the oracle
    rng = np.random.default_rng(3)
    for lo, hi in ((0, 128), (129, 288)):
        ks = [k for k in range(K) if lo <= ns[k] <= hi]
        for _ in range(6):
            k, p = int(rng.choice(ks)), int(rng.integers(P))
            sel = np.sort(rng.choice(T, 4, replace=False))
            pk = pack_trees([wl.trees[p]])
            S = np.ascontiguousarray(wl.blocks[k][sel])
            want = oracle.evaluate(wl.insts[k], pk, S, len(sel), wl.mask, ctxs[k].zp,
                                   need_rows=need_rows_for(pk, wl.mask), threads=THREADS)
            assert_bitwise(pairs[k, p, sel], want, "instance %d rule %d scenarios %s" % (k, p, sel.tolist()))


def test_gp_loop_full_size_replayed_generations(oracle):
    from oracle import need_rows_for
    from sigp.api import FactoryContext
    from sigp.gp_device import DeviceGP
    from sigp.workloads import TEST_SEED, clustered_gp_instance
    inst = clustered_gp_instance()
    ctx = FactoryContext(inst)
    terms = list(NS_TERMINALS) + list(EXTRA)
    mask = term_mask_from(terms)
    tc, ec = G.TreeConfig(terms, list(FUNCTIONS)), G.EvolutionConfig(population_size=1024)
    T, gens, seed = 50, 50, 0
    rows = [gen_durations(inst.durations, T, seed * 10000 + g, mode="gamma", shape=4.0) for g in range(gens)]
    gp = DeviceGP(ctx, tc, ec, seeds=[seed], T=T, term_mask=mask)
    gp.init()
    done = 0
    hist = []
    for g_replay in (17, 38):
        if g_replay > done:
            r = gp.run(np.concatenate(rows[done:g_replay]), generations=g_replay - done, step_last=True)
            hist.append(r.fitness)
        pop = gp.population()[0]
        state = gp.rng_state(0)
        print("replaying generation %d on the host" % g_replay, flush=True)
        r = gp.run(rows[g_replay], generations=1, step_last=True)
        hist.append(r.fitness)
        dev_fit = r.fitness[0, 0]
        pk = pack_trees(pop)
        pairs = oracle.evaluate(inst, pk, rows[g_replay], T, mask, ctx.zp, need_rows=need_rows_for(pk, mask),
                                threads=THREADS)
        assert_bitwise(dev_fit, _means(pairs, T), "fitness of generation %d" % g_replay)
        host = random.Random()
        host.setstate((3, tuple(state), None))
        nxt = G.gp_step(pop, list(dev_fit), host, tc, ec)
        assert gp.prefixes(0) == [to_prefix(t) for t in nxt], "offspring of generation %d" % g_replay
        assert gp.rng_state(0) == tuple(host.getstate()[1])
        done = g_replay + 1
    r = gp.run(np.concatenate(rows[done:]), generations=gens - done)
    hist.append(r.fitness)
    fit = np.concatenate(hist)
    assert fit.shape[:2] == (gens, 1)
    assert np.all(fit == np.round(fit * T) / T)
    # the final generation's best rule (first maximum) on 500 held-out scenarios
    final = gp.population()[0]
    best = final[int(np.argmax(fit[-1, 0]))]
    test = gen_durations(inst.durations, 500, TEST_SEED, mode="gamma", shape=4.0)
    pk = pack_trees([best])
    _, held = ctx.evaluate_means(pk, test, 500, mask)
    want = oracle.evaluate(inst, pk, test, 500, mask, ctx.zp, threads=THREADS)
    assert_bitwise(held, _means(want, 500), "held-out mean")
    print("best rule held-out mean %.3f; best training fitness by generation: %s"
          % (held[0], np.round(fit[:, 0].max(axis=1)[::7], 2).tolist()), flush=True)
