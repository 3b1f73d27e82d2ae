"""The GP loop on the device (sigp_gp_*) against the reference's gp.txt
(populations and fitness of 3 generations x 1,024 trees, two vocabularies)
and against the host GP operators (sigp/gp.py, pinned by the same fixture)
on what gp.txt leaves open: ERC constants, several runs, the pool branch of
sample() for small populations, fitness ties, other elitism and rates, and
the graph-captured loop against stream execution and the host loop."""
import random

import numpy as np
import pytest

from golden_data import gp_cases
from gpu_util import assert_bitwise
from sigp import gp as G
from sigp.instance import load_instance, synthetic_clustered
from sigp.rules import (EXTRA, FUNCS_RMOD, FUNCTIONS, NS_TERMINALS, pack_trees, parse_prefix, term_mask_from,
                        to_prefix)
from sigp.scenarios import gen_durations

pytestmark = pytest.mark.gpu
GP = gp_cases()


def _vocab(kind):
    if kind == "ext29rmod":
        return list(NS_TERMINALS) + list(EXTRA), list(FUNCS_RMOD)
    return list(NS_TERMINALS), list(FUNCTIONS)


def _state(rng):
    return tuple(rng.getstate()[1])


@pytest.mark.parametrize("ci", range(len(GP)), ids=[c["kind"] for c in GP])
def test_device_gp_steps_reproduce_gp_fixture(ci):
    from sigp.api import FactoryContext
    from sigp.gp_device import DeviceGP
    case = GP[ci]
    terms, funcs = _vocab(case["kind"])
    ctx = FactoryContext(load_instance("c101"))
    tc, ec = G.TreeConfig(terms, funcs), G.EvolutionConfig()
    gp = DeviceGP(ctx, tc, ec, seeds=[0], T=1, term_mask=term_mask_from(terms))
    gp.init()
    host = random.Random(0)
    G.ramped_half_and_half(tc, host, ec.population_size)
    for gi, g in enumerate(case["gens"]):
        print("gp", case["kind"], "generation", gi, flush=True)
        assert gp.prefixes(0) == g["pop"], "population of generation %d" % gi
        assert gp.rng_state(0) == _state(host), "generator state before generation %d" % gi
        fit = gp.evaluate(g["sampled"][None, :])
        assert_bitwise(fit[0], g["fits"], "fitness of generation %d" % gi)
        if gi < 2:
            gp.step()   # with the last evaluation's fitness
            # the host generator advances by the same draws
            G.gp_step([parse_prefix(t) for t in g["pop"]], list(g["fits"]), host, tc, ec)


@pytest.mark.parametrize("ci", range(len(GP)), ids=[c["kind"] for c in GP])
@pytest.mark.parametrize("graph", [True, False], ids=["graph", "stream"])
def test_device_gp_loop_reproduces_gp_fixture(ci, graph):
    """Three generations in one call: evaluate, step, evaluate, step, evaluate."""
    from sigp.api import FactoryContext
    from sigp.gp_device import DeviceGP
    case = GP[ci]
    terms, funcs = _vocab(case["kind"])
    ctx = FactoryContext(load_instance("c101"))
    gp = DeviceGP(ctx, G.TreeConfig(terms, funcs), G.EvolutionConfig(), seeds=[0], T=1,
                  term_mask=term_mask_from(terms))
    gp.init()
    rows = np.stack([g["sampled"] for g in case["gens"]])
    print("gp loop", case["kind"], "graph" if graph else "stream", flush=True)
    res = gp.run(rows, generations=3, graph=graph)
    print("3 generations in %.2f ms" % res.ms, flush=True)
    for gi, g in enumerate(case["gens"]):
        assert_bitwise(res.fitness[gi, 0], g["fits"], "fitness of generation %d" % gi)
        # the recorded best rule: first maximum of the generation's fitness
        assert to_prefix(res.best[gi][0]) == g["pop"][int(np.argmax(g["fits"]))]
    assert gp.prefixes(0) == case["gens"][2]["pop"]
    assert np.all(np.diff(res.generation_ms) > 0)


def _host_runs(tc, ec, seeds):
    rngs = [random.Random(s) for s in seeds]
    pops = [G.ramped_half_and_half(tc, r, ec.population_size) for r in rngs]
    return rngs, pops


@pytest.mark.parametrize("pop,elitism,tourn,rates,seeds", [
    (40, 3, 7, (0.8, 0.15), [3, 11]),        # population <= sample()'s set size: pool branch
    (200, 10, 7, (0.8, 0.15), [5, 2 ** 40 + 7]),
    (97, 0, 3, (0.5, 0.4), [0]),
    (64, 70, 5, (0.8, 0.15), [9]),            # elitism above the population
], ids=["pool", "set", "rates", "elite-all"])
def test_device_gp_matches_host_operators(pop, elitism, tourn, rates, seeds):
    """ERC terminals, several runs, ties in the fitness: populations, the
    compiled arrays and generator states equal the host operators' for 4
    generations driven by the same fitness values."""
    from sigp.api import FactoryContext
    from sigp.gp_device import DeviceGP
    terms = ["SCORE", "TA", "NS", "ERC", "REGRET", "TC", "ERC"]
    tc = G.TreeConfig(terms, list(FUNCS_RMOD), max_depth=6, min_depth=1)
    ec = G.EvolutionConfig(population_size=pop, elitism=elitism, tournament_size=tourn,
                           crossover_rate=rates[0], mutation_rate=rates[1], max_depth=7)
    ctx = FactoryContext(load_instance("c101"))
    gp = DeviceGP(ctx, tc, ec, seeds=seeds, T=1, term_mask=term_mask_from([t for t in terms if t != "ERC"]))
    gp.init()
    rngs, pops = _host_runs(tc, ec, seeds)
    fr = np.random.default_rng(17)
    for gen in range(5):
        dev = gp.population()
        for r in range(len(seeds)):
            assert [to_prefix(t) for t in dev[r]] == [to_prefix(t) for t in pops[r]], (gen, r)
            assert gp.rng_state(r) == _state(rngs[r]), (gen, r)
        pk = gp.packed()
        want = pack_trees([t for p in pops for t in p])
        assert np.array_equal(pk.lengths, want.lengths)
        assert np.array_equal(pk.ops, want.ops) and np.array_equal(pk.args, want.args)
        if gen == 4:
            break
        fit = fr.integers(0, 6, size=(len(seeds), pop)).astype(np.float64)   # many ties
        gp.step(fit)
        pops = [G.gp_step(pops[r], list(fit[r]), rngs[r], tc, ec) for r in range(len(seeds))]


def test_device_gp_loop_equals_host_loop():
    """Two runs of 256 rules, 4 scenarios per generation, 5 generations on a
    clustered instance (NS-11 + EXTRA-18): the graph loop, the stream loop and
    the host loop (host operators + device evaluation) agree bit for bit."""
    from sigp.api import FactoryContext
    from sigp.gp_device import DeviceGP
    inst = synthetic_clustered(n=100, m=3, seed=1)
    ctx = FactoryContext(inst)
    terms = list(NS_TERMINALS) + list(EXTRA)
    mask = term_mask_from(terms)
    tc, ec = G.TreeConfig(terms, list(FUNCTIONS)), G.EvolutionConfig(population_size=256)
    seeds, T, gens = [4, 8], 4, 5
    rows = np.concatenate([gen_durations(inst.durations, T, s * 10000 + g, mode="gamma")
                           for g in range(gens) for s in seeds])
    hist = {}
    pops = {}
    for graph in (True, False):
        gp = DeviceGP(ctx, tc, ec, seeds=seeds, T=T, term_mask=mask)
        gp.init()
        print("device loop", "graph" if graph else "stream", flush=True)
        res = gp.run(rows, generations=gens, graph=graph)
        hist[graph] = res.fitness
        print("  %.1f ms" % res.ms, flush=True)
        pops[graph] = [[to_prefix(t) for t in p] for p in gp.population()]
    assert_bitwise(hist[True], hist[False], "graph vs stream")
    assert pops[True] == pops[False]
    rngs, hp = _host_runs(tc, ec, seeds)
    per = len(seeds) * T
    for g in range(gens):
        for r in range(len(seeds)):
            blk = rows[g * per + r * T:g * per + (r + 1) * T]
            _, mean = ctx.evaluate_means(pack_trees(hp[r]), blk, T, mask)
            assert_bitwise(hist[True][g, r], mean, "generation %d run %d" % (g, r))
        if g + 1 < gens:
            hp = [G.gp_step(hp[r], list(hist[True][g, r]), rngs[r], tc, ec) for r in range(len(seeds))]
    assert pops[True] == [[to_prefix(t) for t in p] for p in hp]


def test_device_screened_steps_match_host_restatement():
    """Surrogate-screened steps (no reference counterpart): k = 4 gp_steps on
    the streams Mt19937(seed + s * 2^32), phenotypes on 64 x 8 situations,
    scores from the device network, then the elites and the best-scored
    offspring outside the elites' slots, phenotypic duplicates last -- against
    the same step restated on the host with the host operators,
    sigp_phenotype and sigp_mlp_predict."""
    from sigp.api import FactoryContext
    from sigp.gp_device import SUB_STREAM, DeviceGP
    from sigp.surrogate import Mlp, features, make_situations, phenotype
    inst = load_instance("c101")
    ctx = FactoryContext(inst)
    terms = list(NS_TERMINALS) + list(EXTRA)
    mask = term_mask_from(terms)
    pop, k, seed = 64, 4, 7
    tc = G.TreeConfig(terms, list(FUNCTIONS), max_depth=6)
    ec = G.EvolutionConfig(population_size=pop, elitism=3, max_depth=7)
    gp = DeviceGP(ctx, tc, ec, seeds=[seed], T=2, term_mask=mask, offspring=k * pop)
    net = Mlp(seed=3)
    sit = make_situations(inst, seed=5)
    gp.set_surrogate(net, sit)
    gp.init()
    rngs = [random.Random(seed + s * SUB_STREAM) for s in range(k)]
    host = G.ramped_half_and_half(tc, rngs[0], pop)
    fr = np.random.default_rng(5)
    E = min(ec.elitism, pop)
    for gen in range(4):
        print("screened step", gen, flush=True)
        assert gp.prefixes(0) == [to_prefix(t) for t in host], gen
        for s in range(k):
            assert gp.rng_state(0, s) == _state(rngs[s]), (gen, s)
        if gen == 3:
            break
        fit = fr.integers(0, 9, size=pop).astype(np.float64)
        gp.step(fit[None, :])
        off = []
        for s in range(k):
            off += G.gp_step(host, list(fit), rngs[s], tc, ec)
        ranks = phenotype(ctx, pack_trees(off), mask, sit)
        pred = net.predict(features(ranks))
        seen, dup = set(), []
        for j in range(k * pop):
            h = _fnv1a(ranks[j])
            dup.append(h in seen)
            seen.add(h)
        cand = [j for j in range(k * pop) if j % pop >= E]
        key = lambda j: (dup[j], -float(pred[j]) if pred[j] == pred[j] else float("inf"), j)
        host = off[:E] + [off[j] for j in sorted(cand, key=key)[:pop - E]]
        print("  distinct phenotypes among %d offspring: %d" % (k * pop, k * pop - sum(dup)), flush=True)


def test_device_screened_init_matches_host_restatement():
    """Screened initial population (no reference counterpart): k ramped
    half-and-half populations, one per stream Mt19937(seed + s * 2^32), the
    screen keeping pop of the k * pop trees with no elites -- against the same
    restated on the host; then one screened step from it."""
    from sigp.api import FactoryContext
    from sigp.gp_device import SUB_STREAM, DeviceGP
    from sigp.surrogate import Mlp, features, make_situations, phenotype
    inst = load_instance("r101")
    ctx = FactoryContext(inst)
    terms = list(NS_TERMINALS) + list(EXTRA)
    mask = term_mask_from(terms)
    pop, k, seed = 48, 4, 19
    tc = G.TreeConfig(terms, list(FUNCTIONS), max_depth=6)
    ec = G.EvolutionConfig(population_size=pop, elitism=2, max_depth=7)
    gp = DeviceGP(ctx, tc, ec, seeds=[seed], T=2, term_mask=mask, offspring=k * pop)
    net = Mlp(seed=4)
    sit = make_situations(inst, seed=6)
    gp.set_surrogate(net, sit)
    gp.init(screened=True)
    rngs = [random.Random(seed + s * SUB_STREAM) for s in range(k)]

    def screen(trees, E):
        ranks = phenotype(ctx, pack_trees(trees), mask, sit)
        pred = net.predict(features(ranks))
        seen, dup = set(), []
        for j in range(len(trees)):
            h = _fnv1a(ranks[j])
            dup.append(h in seen)
            seen.add(h)
        cand = [j for j in range(len(trees)) if j % pop >= E]
        key = lambda j: (dup[j], -float(pred[j]) if pred[j] == pred[j] else float("inf"), j)
        return trees[:E] + [trees[j] for j in sorted(cand, key=key)[:pop - E]]

    pool = []
    for s in range(k):
        pool += G.ramped_half_and_half(tc, rngs[s], pop)
    host = screen(pool, 0)
    assert gp.prefixes(0) == [to_prefix(t) for t in host]
    for s in range(k):
        assert gp.rng_state(0, s) == _state(rngs[s]), s
    fit = np.random.default_rng(2).integers(0, 9, size=pop).astype(np.float64)
    gp.step(fit[None, :])
    off = []
    for s in range(k):
        off += G.gp_step(host, list(fit), rngs[s], tc, ec)
    host = screen(off, min(ec.elitism, pop))
    assert gp.prefixes(0) == [to_prefix(t) for t in host]


def _fnv1a(ranks):
    h = 0xCBF29CE484222325
    for v in np.asarray(ranks).ravel():
        h = ((h ^ (int(v) & 0xFF)) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def test_device_gp_loop_in_chunks_equals_one_call():
    """A loop run as 2 + 3 generations (stepping after the first call's last
    generation) equals the same 5 generations in one call."""
    from sigp.api import FactoryContext
    from sigp.gp_device import DeviceGP
    inst = synthetic_clustered(n=100, m=3, seed=1)
    ctx = FactoryContext(inst)
    terms = list(NS_TERMINALS) + list(EXTRA)
    mask = term_mask_from(terms)
    tc, ec = G.TreeConfig(terms, list(FUNCTIONS)), G.EvolutionConfig(population_size=128)
    T, gens = 3, 5
    rows = np.concatenate([gen_durations(inst.durations, T, 6 * 10000 + g, mode="gamma") for g in range(gens)])
    one = DeviceGP(ctx, tc, ec, seeds=[6], T=T, term_mask=mask)
    one.init()
    full = one.run(rows, generations=gens)
    two = DeviceGP(ctx, tc, ec, seeds=[6], T=T, term_mask=mask)
    two.init()
    a = two.run(rows[:2 * T], generations=2, step_last=True)
    b = two.run(rows[2 * T:], generations=3)
    assert_bitwise(np.concatenate([a.fitness, b.fitness]), full.fitness, "chunked vs one call")
    assert [to_prefix(t[0]) for t in a.best + b.best] == [to_prefix(t[0]) for t in full.best]
    assert two.prefixes(0) == one.prefixes(0)
