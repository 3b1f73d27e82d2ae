"""Per-candidate terminal values and the round-robin rollout of the oracle
against the reference's own code, compiled from its sources (oracle/_ref):

  * the reference's pick_best (engine/sim_core.hpp:43-204) with every slot
    masked in leaves the 61 slot values of the last feasible candidate in
    its terms buffer -- all of them, composites and Z slots included, must
    equal the oracle's values for that candidate bit for bit;
  * every scan terminal of engine/sim_terms.hpp (t_ns ... t_scorerank,
    struct_feats_one), called directly for every feasible candidate, must
    equal the oracle's slot value bit for bit;
  * the reference's sim_one_rr (engine/sim_core.hpp:241-274) must equal
    orc_sim_one_rr on random rules and scenarios.

Decision states are partial routes reached by random feasible steps from the
depot (with sampled service times, so the clock is off the expected path, and
with vehicle changes) plus states drawn at random (visited sets of every
density, any current POI, any clock), on c101, r101, rc101, pr01 and a
synthetic 288-customer clustered + uniform instance.
"""
import dataclasses
import random
import zlib

import numpy as np
import pytest

from golden_data import bits
from sigp.instance import Instance, load_instance, synthetic_mixed, synthetic_uniform
from sigp.rules import EXTRA, FUNCS_RMOD, NS_TERMINALS, SLOT_NAMES, pack_trees, term_mask_from
from sigp.scenarios import gen_durations
from sigp.workloads import random_rules

# slots computed by one terminal function of their own (engine/sim_core.hpp:99-197)
SCAN_SLOTS = [10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 30, 41, 44, 47, 48, 49, 50,
              53, 54, 55, 56, 57, 59, 60]
FULL = np.ones(61, dtype=np.int64)

INSTANCES = {"c101": 200, "r101": 200, "rc101": 200, "pr01": 300, "syn288": 60, "synvar": 200}


def _varied_durations(n=80, m=3, seed=11):
    """A uniform synthetic instance whose expected service times are
    non-integer and differ per customer (U[1, 30]) on a short horizon: the
    benchmark instances have one service time each (c101, r101, rc101, the
    synthetic ones) or a horizon long enough to saturate the Gaussian tails
    (pr01), which would hide a slip between a candidate's and a neighbour's
    duration."""
    base = synthetic_uniform(n=n, m=m, seed=seed)
    rng = random.Random(seed)
    pois = [base.pois[0]] + [dataclasses.replace(p, duration=round(rng.uniform(1.0, 30.0), 3))
                             for p in base.pois[1:]]
    return Instance("synvar", pois, base.tmax, m)


def _instance(name):
    if name == "syn288":
        return synthetic_mixed(n=288, m=4, seed=1)
    if name == "synvar":
        return _varied_durations()
    return load_instance(name)


def _walk_states(oracle, inst, rng, count, zp):
    """States along random feasible routes: each step takes a random
    feasible customer and advances the clock by a sampled service time
    (expected x U[0.5, 1.6]); a vehicle ends when nothing is feasible."""
    out = []
    const = (np.array([-2], dtype=np.int32), np.zeros(2, dtype=np.int32))
    mask1 = np.zeros(61, dtype=np.int64)
    while len(out) < count:
        vis = np.zeros(inst.n, dtype=np.uint8)
        for v in range(inst.m):
            rem, cur, t = inst.m - v - 1, 0, 0.0
            while len(out) < count:
                out.append((cur, t, rem, vis.copy()))
                _, pri = oracle.pick_best(inst, cur, t, rem, const[0], const[1], vis, mask1, zp)
                feas = np.flatnonzero(np.isfinite(pri))
                if feas.size == 0:
                    break
                b = int(feas[rng.randrange(feas.size)])
                arr = t + float(np.asarray(inst.dist).reshape(-1)[cur * (inst.n + 1) + b + 1])
                start = max(arr, float(inst.open_t[b]))
                t = start + float(inst.durations[b]) * rng.uniform(0.5, 1.6)
                cur = b + 1
                vis[b] = 1
    return out


def _random_states(inst, rng, count):
    out = []
    for _ in range(count):
        dens = rng.choice([0.0, 0.05, 0.3, 0.6, 0.9])
        vis = np.array([rng.random() < dens for _ in range(inst.n)], dtype=np.uint8)
        cur = rng.randrange(inst.n + 1)
        if cur > 0:
            vis[cur - 1] = 1
        t = rng.uniform(0.0, float(inst.tmax)) if rng.random() < 0.8 else 0.0
        out.append((cur, t, rng.randrange(max(inst.m, 1)), vis))
    return out


@pytest.fixture(scope="module", params=sorted(INSTANCES))
def states(request, oracle):
    name = request.param
    inst = _instance(name)
    zp = oracle.zstats(inst)
    rng = random.Random(zlib.crc32(name.encode()))
    k = INSTANCES[name]
    sts = _walk_states(oracle, inst, rng, k, zp) + _random_states(inst, rng, k)
    return name, inst, zp, sts


def test_last_feasible_candidate_all_slots(oracle, reference, states):
    """All 61 slots, as the reference's own pick_best fills them."""
    name, inst, zp, sts = states
    compared = 0
    for cur, t, rem, vis in sts:
        vals, feas = oracle.terms_all(inst, cur, t, rem, vis, FULL, zp)
        b, ref_terms = reference.pick_best_terms(inst, cur, t, rem, vis, FULL, zp)
        idx = np.flatnonzero(feas)
        assert b == (int(idx[0]) if idx.size else -1), (name, cur, t)
        if idx.size == 0:
            continue
        last = int(idx[-1])
        diff = np.flatnonzero(bits(vals[last]) != bits(ref_terms))
        assert diff.size == 0, (name, cur, t, rem, last, diff.tolist(), vals[last][diff], ref_terms[diff])
        compared += 1
    assert compared >= len(sts) // 3


def test_scan_terminals_every_candidate(oracle, reference, states):
    """Every scan terminal, every feasible candidate."""
    name, inst, zp, sts = states
    n_vals = 0
    for cur, t, rem, vis in sts:
        vals, feas = oracle.terms_all(inst, cur, t, rem, vis, FULL, zp)
        cands = np.flatnonzero(feas)
        if cands.size == 0:
            continue
        for slot in SCAN_SLOTS:
            want = reference.scan_terminal(inst, slot, cur, t, vis, cands)
            got = vals[cands, slot]
            bad = np.flatnonzero(bits(got) != bits(want))
            assert bad.size == 0, (name, slot, cur, t, cands[bad].tolist(), got[bad], want[bad])
            n_vals += cands.size
    assert n_vals > 0


def test_scan_states_are_varied(states):
    """The states reach late clocks, dense visited sets and every vehicle."""
    name, inst, zp, sts = states
    ts = np.array([s[1] for s in sts])
    dens = np.array([s[3].mean() for s in sts])
    assert ts.max() > 0.5 * inst.tmax and (ts == 0).any()
    assert dens.max() > 0.5 and (dens == 0).any()
    assert len({s[2] for s in sts}) == max(inst.m, 1)


@pytest.mark.parametrize("name,terms,funcs,m", [
    ("c101", NS_TERMINALS + EXTRA, None, 3), ("r101", SLOT_NAMES, None, 2),
    ("pr01", NS_TERMINALS + EXTRA, FUNCS_RMOD, 4), ("rc101", EXTRA, None, 1),
    ("syn288", EXTRA, None, 4), ("synvar", SLOT_NAMES, FUNCS_RMOD, 3)])
def test_sim_one_rr_equals_reference(oracle, reference, name, terms, funcs, m):
    inst = _instance(name).with_m(m)
    kw = {} if funcs is None else {"functions": funcs}
    trees = random_rules(12 if name == "syn288" else 24, terms, seed=5, min_depth=1, max_depth=5, **kw)
    pk = pack_trees(trees)
    S = gen_durations(inst.durations, 4, 77, mode="gamma")
    mask = term_mask_from(terms)
    zp = oracle.zstats(inst)
    got, want = [], []
    for p in range(pk.n_rules):
        o, L = int(pk.offsets[p]), int(pk.lengths[p])
        ops, args = pk.ops[o:o + L], pk.args[2 * o:2 * (o + L)]
        for s in range(S.shape[0]):
            got.append(oracle.sim_one_rr(inst, ops, args, S[s], mask, zp))
            want.append(reference.sim_one_rr(inst, ops, args, S[s], mask, zp))
    assert np.array_equal(bits(got), bits(want))
    # the round-robin arm is a different protocol: it does not reproduce V0 everywhere
    v0 = oracle.evaluate(inst, pk, S, S.shape[0], mask, zp, threads=8)
    assert np.all(np.asarray(got) >= 0) and np.all(np.asarray(got) == np.round(got))
    if m > 1:
        assert not np.array_equal(np.asarray(got), v0)
