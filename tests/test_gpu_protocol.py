"""The round-robin protocol on the device (SIGP_PROTOCOL_ROUND_ROBIN): the
reference's sim_one_rr (engine/sim_core.hpp:241-274), the V2 ablation arm it
runs on the CPU only.  Every mapping, with and without the scenario order,
single- and multi-instance batches and the allocation-free entry point give
the oracle's orc_sim_one_rr bit for bit (the oracle's function is itself
pinned to the reference's sim_one_rr, compiled from its sources, by
tests/test_oracle_terminals.py)."""
import os

import numpy as np
import pytest

from gpu_util import assert_bitwise
from sigp.instance import load_instance, synthetic_mixed
from sigp.rules import EXTRA, FUNCS_RMOD, NS_TERMINALS, pack_trees, term_mask_from
from sigp.scenarios import gen_durations
from sigp.workloads import random_rules

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
MAPS = [(1, w) for w in (1, 2, 4)] + [(2, w) for w in (1, 2)] + [(3, 1)]   # (mapping, waves)
FULL61 = np.ones(61, dtype=np.int64)


def rr_oracle(oracle, inst, pk, S, mask, zp):
    """out[p * T + s] = orc_sim_one_rr(rule p, scenario row s)."""
    return oracle.evaluate_rr(inst, pk, S, S.shape[0], mask, zp, threads=THREADS)


CASES = [("c101", 3, "ns18", None), ("rc101", 4, "full61", None), ("pr01", 2, "ns18", FUNCS_RMOD),
         ("syn288", 4, "ns18", None)]


@pytest.mark.parametrize("name,m,vocab,funcs", CASES)
def test_round_robin_equals_oracle_every_mapping(oracle, name, m, vocab, funcs):
    from sigp.api import FactoryContext
    base = synthetic_mixed(n=288, m=4, seed=3) if name == "syn288" else load_instance(name)
    inst = base.with_m(m)
    terms = NS_TERMINALS + EXTRA
    kw = {} if funcs is None else {"functions": funcs}
    trees = random_rules(8 if name == "syn288" else 20, terms, seed=41, min_depth=1, max_depth=5, **kw)
    pk = pack_trees(trees)
    mask = FULL61 if vocab == "full61" else term_mask_from(terms)
    T = 70   # two blocks of scenario lanes, the second partial
    S = gen_durations(inst.durations, T, 123, mode="gamma")
    ctx = FactoryContext(inst)
    want = rr_oracle(oracle, inst, pk, S, mask, ctx.zp)
    v0 = oracle.evaluate(inst, pk, S, T, mask, ctx.zp, threads=THREADS)
    assert not np.array_equal(want, v0)   # m > 1: a different protocol
    ctx.protocol(True)
    for mp, w in MAPS:
        print("round robin", name, "m", m, "mapping", mp, "waves", w, flush=True)
        ctx.configure(w, mp)
        assert_bitwise(ctx.evaluate(pk, S, T, mask), want, "%s mapping %d waves %d" % (name, mp, w))
    ctx.protocol(False)
    ctx.configure(0, 0)
    assert_bitwise(ctx.evaluate(pk, S, T, mask), v0, "%s back to vehicle by vehicle" % name)


def test_round_robin_one_vehicle_is_sim_one(oracle):
    """With m = 1 the two protocols are the same simulation."""
    from sigp.api import FactoryContext
    inst = load_instance("r101").with_m(1)
    trees = random_rules(24, NS_TERMINALS + EXTRA, seed=2, min_depth=1, max_depth=5)
    pk = pack_trees(trees)
    mask = term_mask_from(NS_TERMINALS + EXTRA)
    S = gen_durations(inst.durations, 64, 9, mode="gamma")
    ctx = FactoryContext(inst)
    v0 = ctx.evaluate(pk, S, 64, mask)
    ctx.protocol(True)
    assert_bitwise(ctx.evaluate(pk, S, 64, mask), v0, "m = 1")


def test_round_robin_batches_grouped_multi_noalloc(oracle):
    """Batch-level switch with the scenario order on, grouped masks, a
    multi-instance batch and sigp_evaluate_noalloc."""
    import torch
    from sigp.api import Batch, FactoryContext, batch_multi
    import ctypes as C
    from sigp._lib import check, lib
    terms = NS_TERMINALS + EXTRA
    mask = term_mask_from(terms)
    insts = [load_instance("c201").with_m(3), load_instance("r102").with_m(2),
             synthetic_mixed(n=150, m=4, seed=8)]
    trees = random_rules(16, terms, seed=77, min_depth=1, max_depth=5)
    pk = pack_trees(trees)
    T = 96
    ctxs = [FactoryContext(i) for i in insts]
    rows = [gen_durations(i.durations, T, 500 + k, mode="gamma") for k, i in enumerate(insts)]
    want = [rr_oracle(oracle, i, pk, S, mask, c.zp) for i, S, c in zip(insts, rows, ctxs)]

    # single-instance batch, switched to round robin after creation, scenario order on
    S0 = torch.tensor(rows[0], device="cuda")
    out = torch.zeros(pk.n_rules * T, dtype=torch.float64, device="cuda")
    bt = Batch(ctxs[0], pk, T, pk.n_rules, mask, pk.n_rules)
    assert bt.order_scenarios(True)
    bt.protocol(True)
    bt.run(S0.data_ptr(), T, out.data_ptr())
    torch.cuda.synchronize()
    assert_bitwise(out.cpu().numpy(), want[0], "batch switched to round robin")
    bt.protocol(False)
    bt.run(S0.data_ptr(), T, out.data_ptr())
    torch.cuda.synchronize()
    v0 = oracle.evaluate(insts[0], pk, rows[0], T, mask, ctxs[0].zp, threads=THREADS)
    assert_bitwise(out.cpu().numpy(), v0, "batch switched back")

    # grouped masks: rule p reads rows (p / 4) * T' + s and mask row p / 8
    masks = np.stack([term_mask_from(NS_TERMINALS), term_mask_from(terms)])
    Tg = 5
    Sg = gen_durations(insts[1].durations, 4 * Tg, 31, mode="gamma")
    ctxs[1].protocol(True)
    got = ctxs[1].evaluate_grouped_masks(pk, Sg, Tg, 4, 8, masks)
    wg = oracle.evaluate_rr(insts[1], pk, Sg, Tg, masks, ctxs[1].zp, group_size=4, mask_group=8, threads=THREADS)
    assert_bitwise(got, wg, "grouped masks")

    # multi-instance batch: the instances must agree on the protocol
    with pytest.raises(Exception):
        batch_multi(ctxs, pk, T, mask)
    for c in ctxs:
        c.protocol(True)
    bm = batch_multi(ctxs, pk, T, mask)
    Sm = torch.tensor(np.concatenate([r.ravel() for r in rows]), device="cuda")
    outm = torch.zeros(len(insts) * pk.n_rules * T, dtype=torch.float64, device="cuda")
    bm.run(Sm.data_ptr(), len(insts) * T, outm.data_ptr())
    torch.cuda.synchronize()
    assert_bitwise(outm.cpu().numpy(), np.concatenate(want), "multi-instance round robin")

    # the allocation-free entry point takes the instance's protocol too
    from sigp.api import _masks, _rules
    keep = []
    r = _rules(pk, keep)
    mk = _masks(mask)
    out2 = torch.zeros(pk.n_rules * T, dtype=torch.float64, device="cuda")
    check(lib().sigp_evaluate_noalloc(ctxs[0]._h, C.byref(r), C.c_void_p(S0.data_ptr()), T, T, pk.n_rules,
                                      mk.ctypes.data_as(C.c_void_p), mk.shape[0], pk.n_rules,
                                      C.c_void_p(out2.data_ptr()), None, None))
    torch.cuda.synchronize()
    assert_bitwise(out2.cpu().numpy(), want[0], "noalloc round robin")


def test_round_robin_rejects_many_vehicles():
    from sigp.api import FactoryContext
    from sigp._lib import SigpError
    ctx = FactoryContext(load_instance("c101").with_m(9))
    with pytest.raises(SigpError):
        ctx.protocol(True)
