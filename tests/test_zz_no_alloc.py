"""sigp_evaluate_noalloc (device buffers, fixed per-GPU scratch, no device
allocation) through the C ABI against sigp_evaluate, bit for bit: two calls
in a row on one instance, the second with a smaller rule set and fewer
scenarios (so the scratch array holds the first call's larger contents when
the second runs), rewards and means; then a rule set too large for the
scratch array, rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import assert_bitwise
from sigp.instance import load_instance, synthetic_uniform
from sigp.rules import EXTRA, FUNCS_RMOD, NS_TERMINALS, pack_trees, term_mask_from
from sigp.scenarios import gen_durations
from sigp.workloads import random_rules

pytestmark = pytest.mark.gpu


def _noalloc(ctx, pk, S, T, masks, group_size=None, mask_group=None):
    import torch
    from sigp._lib import check, lib
    from sigp.api import _masks, _rules
    keep = []
    R = _rules(pk, keep)
    m = _masks(masks)
    P = pk.n_rules
    dev = torch.device("cuda:0")
    sd = torch.from_numpy(np.ascontiguousarray(S, dtype=np.float64)).to(dev)
    pairs = torch.full((P * T,), float("nan"), dtype=torch.float64, device=dev)
    means = torch.full((P,), float("nan"), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    check(lib().sigp_evaluate_noalloc(ctx._h, C.byref(R), C.c_void_p(sd.data_ptr()), S.shape[0], T,
                                      group_size or P, m.ctypes.data_as(C.c_void_p), m.shape[0],
                                      mask_group or P, C.c_void_p(pairs.data_ptr()),
                                      C.c_void_p(means.data_ptr()), C.c_void_p(stream.cuda_stream)))
    return pairs.cpu().numpy(), means.cpu().numpy()


def test_noalloc_equals_evaluate_two_calls():
    from sigp.api import FactoryContext
    inst = synthetic_uniform(n=100, m=2, seed=1)
    ctx = FactoryContext(inst)
    mask = term_mask_from(EXTRA)
    # call 1: 320 rules (rmod included) x 128 scenarios -- scenario lanes
    trees = random_rules(320, EXTRA, seed=3, functions=FUNCS_RMOD)
    pk = pack_trees(trees)
    S = gen_durations(inst.durations, 128, 1_000_000, mode="gamma")
    print("noalloc call 1: P=%d T=%d" % (pk.n_rules, 128), flush=True)
    got_p, got_m = _noalloc(ctx, pk, S, 128, mask)
    want_p, want_m = ctx.evaluate_means(pk, S, 128, mask)
    assert_bitwise(got_p, want_p, "noalloc call 1 rewards")
    assert_bitwise(got_m, want_m, "noalloc call 1 means")
    # call 2, smaller: 40 rules x 20 scenarios -- candidate lanes
    trees2 = random_rules(40, EXTRA, seed=4)
    pk2 = pack_trees(trees2)
    S2 = gen_durations(inst.durations, 20, 7, mode="gamma")
    print("noalloc call 2: P=%d T=%d" % (pk2.n_rules, 20), flush=True)
    got_p2, got_m2 = _noalloc(ctx, pk2, S2, 20, mask)
    want_p2, want_m2 = ctx.evaluate_means(pk2, S2, 20, mask)
    assert_bitwise(got_p2, want_p2, "noalloc call 2 rewards")
    assert_bitwise(got_m2, want_m2, "noalloc call 2 means")


def test_noalloc_grouped_masks_equals_evaluate():
    """evaluate_grouped_masks index maths (engine/factory.cuh:98-102) through
    the allocation-free entry point: rule p reads scenario group
    p / group_size and mask row p / mask_group."""
    from sigp.api import FactoryContext
    inst = load_instance("r105")
    ctx = FactoryContext(inst)
    trees = random_rules(24, NS_TERMINALS + EXTRA, seed=21)
    pk = pack_trees(trees)
    masks = np.stack([term_mask_from(NS_TERMINALS), term_mask_from(NS_TERMINALS + EXTRA),
                      term_mask_from(NS_TERMINALS + ["REGRET", "ATRISK"])])
    T, gs, mg = 6, 4, 8
    S = gen_durations(inst.durations, (24 // gs) * T, 11)
    print("noalloc grouped: P=24 T=%d group_size=%d mask_group=%d" % (T, gs, mg), flush=True)
    got_p, _ = _noalloc(ctx, pk, S, T, masks, group_size=gs, mask_group=mg)
    want = ctx.evaluate_grouped_masks(pk, S, T, gs, mg, masks)
    assert_bitwise(got_p, want, "noalloc grouped masks")


def test_noalloc_rejects_oversize_rule_sets():
    """Rule sets whose compiled block exceeds SIGP_NOALLOC_MAX_BYTES fail
    with SIGP_INVALID_ARG on the host, before the device is touched."""
    from sigp._lib import SigpError
    from sigp.api import FactoryContext
    inst = synthetic_uniform(n=20, m=1, seed=2)
    ctx = FactoryContext(inst)
    trees = random_rules(16000, EXTRA, seed=5, min_depth=6, max_depth=6)   # up to 127-node rules
    pk = pack_trees(trees)
    assert pk.n_nodes * 12 > (8 << 20)                # the postfix arrays alone exceed the array
    S = gen_durations(inst.durations, 1, 3)
    with pytest.raises(SigpError) as ei:
        _noalloc(ctx, pk, S, 1, term_mask_from(EXTRA))
    assert ei.value.status == 1 and "SIGP_NOALLOC_MAX_BYTES" in str(ei.value)
