"""Checks that pin the oracle beyond the golden fixtures (CPU only):
the reference's own simulator on random rules and instances, closed forms,
monotone terminal families, brute force on tiny instances, and the
pruning rule the GPU evaluator uses."""
import itertools
import math

import numpy as np
import pytest

from golden_data import bits
from oracle import need_rows_for
from sigp.instance import POI, Instance, load_instance, synthetic_clustered, synthetic_uniform
from sigp.rules import (EXTRA, FUNCS_RMOD, NS_TERMINALS, SLOT_NAMES, pack_trees, parse_prefix,
                        term_mask_from)
from sigp.scenarios import gen_durations
from sigp.workloads import random_rules


@pytest.mark.parametrize("name,terms,funcs", [
    ("c103", SLOT_NAMES, None), ("pr03", NS_TERMINALS + EXTRA, FUNCS_RMOD),
    ("syn", EXTRA, None)])
def test_oracle_equals_reference_build(oracle, reference, name, terms, funcs):
    inst = synthetic_uniform(n=60, m=2, seed=4) if name == "syn" else load_instance(name)
    kw = {} if funcs is None else {"functions": funcs}
    trees = random_rules(24, terms, seed=17, min_depth=2, max_depth=5, **kw)
    pk = pack_trees(trees)
    S = gen_durations(inst.durations, 12, 31, mode="gamma")
    mask = term_mask_from(terms)
    zp = oracle.zstats(inst)
    assert np.array_equal(bits(zp), bits(reference.zp_of(inst)))
    got = oracle.evaluate(inst, pk, S, 12, mask, zp, threads=8)
    want = reference.evaluate(inst, pk, S, 12, mask, zp, threads=8)
    assert np.array_equal(bits(got), bits(want))


def test_pruned_equals_full(oracle):
    """Computing only the live slots a rule reads changes no reward."""
    inst = load_instance("rc102")
    terms = NS_TERMINALS + EXTRA + ["ZWLOSS", "ZREACH", "SKIPLOSS", "NSA05"]
    trees = random_rules(40, terms, seed=8, min_depth=1, max_depth=4)
    pk = pack_trees(trees)
    mask = term_mask_from(NS_TERMINALS + EXTRA + ["ZWLOSS", "ZATRISK"])
    S = gen_durations(inst.durations, 10, 4)
    zp = oracle.zstats(inst)
    full = oracle.evaluate(inst, pk, S, 10, mask, zp, threads=8)
    pruned = oracle.evaluate(inst, pk, S, 10, mask, zp, need_rows=need_rows_for(pk, mask), threads=8)
    assert np.array_equal(bits(full), bits(pruned))


def _state(inst, visited_idx=(), current=0, ctime=0.0):
    vis = np.zeros(inst.n, dtype=np.uint8)
    vis[list(visited_idx)] = 1
    return vis, current, ctime


def test_closed_form_slots(oracle):
    inst = load_instance("r101")
    zp = oracle.zstats(inst)
    vis, cur, t = _state(inst, (3, 9), current=5, ctime=20.0)
    i = 0   # customer 1 is feasible from POI 5 at t=20 in r101
    D = inst.dist
    ta = D[cur, i + 1]
    T = lambda s: oracle.terminal(inst, s, i, cur, t, 1, vis, zp)
    assert T(0) == inst.scores[i] and T(1) == inst.durations[i]
    assert T(2) == inst.open_t[i] - t and T(3) == inst.close_t[i] - t and T(4) == ta
    assert T(5) == inst.dist_to_depot[i]
    assert T(6) == max(inst.open_t[i] - t, ta) and T(7) == max(inst.open_t[i] - t, ta) + inst.durations[i]
    assert T(8) == (inst.close_t[i] - t) - ta and T(9) == 1 * inst.tmax + (inst.tmax - t)
    assert T(24) == t / inst.tmax and T(25) == 2 / inst.n and T(27) == 1.0
    assert [T(s) for s in range(36, 41)] == [0.1, 0.2, 0.3, 0.4, 0.5]
    assert T(31) == (T(3) - zp[0]) / zp[5] and T(35) == (T(5) - zp[4]) / zp[9]
    assert T(42) == (ta + inst.durations[i]) / max(inst.tmax - t, 1.0)
    assert T(58) == D[cur, 0] + inst.dist_to_depot[i] - ta


def test_rand_slot_is_the_lcg(oracle):
    inst = load_instance("c101")
    zp = oracle.zstats(inst)
    vis = np.zeros(inst.n, dtype=np.uint8)
    for i in range(5):
        x = ((i + 1) * 1103515245 + 12345) & 0x7FFFFFFF
        x = (x * 1103515245 + 12345) & 0x7FFFFFFF
        want = (x % 100000) / 1000.0
        assert oracle.terminal(inst, 16, i, 0, 0.0, 2, vis, zp) == want


def test_monotone_families(oracle):
    """REACH1 <= REACH3 <= REACH5 (prefix sums of one greedy path),
    MAXN <= NS, NSROB <= NS (stricter windows, later start)."""
    inst = synthetic_clustered(n=60, m=3, seed=2)
    zp = oracle.zstats(inst)
    rng = np.random.default_rng(3)
    for _ in range(30):
        vis = (rng.random(inst.n) < 0.2).astype(np.uint8)
        cur = int(rng.integers(0, inst.n + 1))
        t = float(rng.uniform(0, 80))
        for i in range(inst.n):
            if vis[i]:
                continue
            r1, r3, r5 = (oracle.terminal(inst, s, i, cur, t, 1, vis, zp) for s in (21, 22, 23))
            if r1 == -math.inf:
                continue
            assert r1 <= r3 <= r5
            ns = oracle.terminal(inst, 10, i, cur, t, 1, vis, zp)
            assert oracle.terminal(inst, 13, i, cur, t, 1, vis, zp) <= ns
            assert oracle.terminal(inst, 12, i, cur, t, 1, vis, zp) <= ns


def _tiny(n=5, seed=0):
    rng = np.random.default_rng(seed)
    pois = [POI(0, 50.0, 50.0, 0.0, 0.0, 300.0, 0.0)]
    for k in range(1, n + 1):
        pois.append(POI(k, float(rng.integers(0, 100)), float(rng.integers(0, 100)),
                        float(rng.integers(1, 30)), 0.0, float(rng.uniform(60, 300)), 10.0))
    return Instance("tiny", pois, 300.0, 1)


def _feasible(inst, q, at, t):
    a = t + inst.dist[at, q + 1]
    if a > inst.close_t[q]:
        return False, 0.0
    st = max(a, inst.open_t[q])
    return st + inst.durations[q] + inst.dist_to_depot[q] <= inst.tmax, st


def test_reach_k_brute_force(oracle):
    """REACH-K by direct enumeration of the greedy path on tiny instances."""
    for seed in range(6):
        inst = _tiny(5, seed)
        zp = oracle.zstats(inst)
        vis = np.zeros(inst.n, dtype=np.uint8)
        for c in range(inst.n):
            ok, st = _feasible(inst, c, 0, 0.0)
            if not ok:
                continue
            for K, slot in ((1, 21), (3, 22), (5, 23)):
                used = {c}; cur = c + 1; t = st + inst.durations[c]; tot = 0.0
                for _ in range(K):
                    best, bv = -1, -1.0
                    for q in range(inst.n):
                        if q in used:
                            continue
                        okq, stq = _feasible(inst, q, cur, t)
                        if okq:
                            v = inst.scores[q] / max(inst.dist[cur, q + 1], 1e-10)
                            if v > bv:
                                bv, best = v, q
                    if best < 0:
                        break
                    _, stb = _feasible(inst, best, cur, t)
                    t = stb + inst.durations[best]; cur = best + 1; tot += inst.scores[best]
                    used.add(best)
                assert oracle.terminal(inst, slot, c, 0, 0.0, 0, vis, zp) == tot


def test_route_brute_force_single_vehicle(oracle):
    """A rule that prefers the lowest index (priority -index via DUR/TA is
    instance-dependent), checked with SCORE: the route must be the greedy
    max-score feasible sequence, ties to the lowest index."""
    for seed in range(5):
        inst = _tiny(6, seed)
        zp = oracle.zstats(inst)
        S = gen_durations(inst.durations, 3, seed)
        ops, args = [-1], [0, 0]
        got = oracle.sim_batch(inst, ops, args, S, 3, term_mask_from(NS_TERMINALS), zp)
        for s in range(3):
            vis = set(); cur = 0; t = 0.0; tot = 0.0
            while True:
                best, bp = -1, -1e300
                for i in range(inst.n):
                    if i in vis:
                        continue
                    ok, st = _feasible(inst, i, cur, t)
                    if ok and inst.scores[i] > bp:
                        bp, best = inst.scores[i], i
                if best < 0:
                    break
                a = t + inst.dist[cur, best + 1]
                t = max(a, inst.open_t[best]) + S[s, best]
                cur = best + 1; tot += inst.scores[best]; vis.add(best)
            assert got[s] == tot
