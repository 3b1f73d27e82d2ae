"""The reference's GP trace (s3_generation/gen_trace.cu) with the GP loop on
the device.

gen_trace trains GP for each of n seeds (population ramped_half_and_half from
Mt19937(seed), vocabulary NS-11 plus the extra terminals) with the generation
loop on the host, and records per seed and generation that generation's best
rule (its train fitness rounded to 1 decimal, size and extra terminals), then
each seed's final rule (the first maximum on one more training draw) with its
gain on 500 held-out scenarios against the NS-GP mean of the baseline table.
trace() runs the n loops side by side on the device (sigp_gp_*, the same
populations as the host loop) and returns the same record, field for field,
as the tool's JSON file holds it.
"""
from __future__ import annotations

import math
from typing import Optional, Sequence

import numpy as np

from . import gp as G
from .api import FactoryContext
from .gp_device import DeviceGP, tokens_to_tree
from .rules import FUNCTIONS, NS_TERMINALS, Node, pack_trees, term_mask_from, to_prefix, tree_size
from .scenarios import gen_durations

TEST_SEED, N_TEST = 1000000, 500                          # s3_generation/gen_trace.cu:19-20
OPS = {"+", "-", "*", "/", "min", "max", "rmod"}           # :21


def round_half_even(x: float, k: int) -> float:
    """engine/jwrite.hpp:10-20, operation for operation."""
    m = math.pow(10.0, k)
    v = x * m
    f = math.floor(v)
    d = v - f
    if d > 0.5:
        r = f + 1
    elif d < 0.5:
        r = f
    else:
        r = f if math.fmod(f, 2.0) == 0.0 else f + 1
    return r / m


def _is_num(t: str) -> bool:
    u = t.replace(".", "")
    return u != "" and all("0" <= c <= "9" for c in u)


def extras_of(rule: Node, base: set) -> list:
    """Sorted distinct tokens of the rule that are neither operators, base
    terminals nor numbers (:36-48)."""
    toks = set(to_prefix(rule).replace("(", " ").replace(")", " ").split())
    return [x for x in sorted(toks) if x not in OPS and x not in base and not _is_num(x)]


def trace(ctx: FactoryContext, name: str, extra: Sequence[str], km: float, n_seeds: int = 30,
          evo: Optional[G.EvolutionConfig] = None) -> dict:
    """gen_trace <name> --terms <extra> --seeds <n_seeds>, as the dict its
    results/trace_<name>_<tag>.json holds (:60-186); km: the instance's
    baseline_nsgp_mean30."""
    ec = evo or G.EvolutionConfig()
    vocab = list(NS_TERMINALS) + list(extra)
    base = set(NS_TERMINALS)
    tc = G.TreeConfig(vocab, list(FUNCTIONS))
    mask = term_mask_from(vocab)
    T, gens = ec.train_samples, ec.generations
    seeds = list(range(n_seeds))
    dur = ctx.inst.durations
    gp = DeviceGP(ctx, tc, ec, seeds=seeds, T=T, term_mask=mask)
    try:
        gp.init()
        rows = np.concatenate([gen_durations(dur, T, s * 10000 + g) for g in range(gens) for s in seeds])
        res = gp.run(rows, generations=gens)
        fit = gp.evaluate(np.concatenate([gen_durations(dur, T, s * 10000 + gens) for s in seeds]))
        tok, ln = gp.tokens()
    finally:
        gp.close()
    tr = [[{"g": g, "fit": round_half_even(float(res.fitness[g, s].max()), 1),
            "sz": tree_size(res.best[g][s]), "ex": extras_of(res.best[g][s], base)}
           for g in range(gens)] for s in seeds]
    best = [int(np.argmax(fit[s])) for s in seeds]         # first maximum (:150-152)
    rules = [tokens_to_tree(tok[s, b, :ln[s, b]], tc.terminals, tc.functions) for s, b in zip(seeds, best)]
    test = gen_durations(dur, N_TEST, TEST_SEED)
    _, means = ctx.evaluate_means(pack_trees(rules), test, N_TEST, mask)
    gains = [100.0 * (float(m) - km) / km for m in means]
    finals = [{"seed": s, "gain": round_half_even(g, 2), "sz": tree_size(r), "ex": extras_of(r, base),
               "rule": to_prefix(r)} for s, (g, r) in enumerate(zip(gains, rules))]
    mean_g = 0.0
    for g in gains:
        mean_g += g
    mean_g /= len(gains)
    var = 0.0
    for g in gains:
        var += (g - mean_g) * (g - mean_g)
    std_g = math.sqrt(var / (len(gains) - 1)) if len(gains) > 1 else 0.0
    return {"instance": name, "vocab_extra": list(extra), "mean_gain": round_half_even(mean_g, 2),
            "std": round_half_even(std_g, 2), "seeds_positive": "%d/%d" % (sum(g > 0 for g in gains), len(gains)),
            "best_seed_gain": round_half_even(max(gains), 2), "finals": finals, "trace": tr}
