"""The GA subset search's proxy fitness with its inner GP on the device.

s4_selection/ga.cu:109-198 (batch_score) scores a genome -- a subset of the
terminal universe, added to NS-11 -- by training GP from scratch for each of
`proxy` seeds: population ramped_half_and_half from Mt19937(seed), `pgens`
evaluations on gen_durations(seed * 10000 + gen) (each but the last followed
by gp_step, on the host), one more evaluation on seed * 10000 + pgens whose
first maximum is the seed's final rule, and that rule's mean over the train
block gen_durations(N_TRAIN = 64, TRAIN_SEED = 2,000,000).  The proxy is the
list of those per-seed means, which the tool caches as hexfloats.

proxy_scores runs the same computation with every seed's GP loop on the
device side by side (sigp_gp_*, the same populations as the host loop), so
the values are the tool's bit for bit.
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from . import gp as G
from .api import FactoryContext
from .gp_device import DeviceGP, tokens_to_tree
from .rules import FUNCTIONS, NS_TERMINALS, pack_trees, term_mask_from
from .scenarios import gen_durations

TRAIN_SEED, N_TRAIN = 2000000, 64        # s4_selection/ga.cu:31-32


def vocab_of(genome: int, universe: Sequence[str]) -> list:
    """NS-11 plus the universe's terminals whose bit is set
    (s4_selection/ga.cu:89-94)."""
    return list(NS_TERMINALS) + [t for b, t in enumerate(universe) if genome >> b & 1]


def proxy_scores(ctx: FactoryContext, universe: Sequence[str], genome: int, proxy: int = 16, pgens: int = 51,
                 evo: Optional[G.EvolutionConfig] = None) -> np.ndarray:
    """Per-seed train-block means of the genome's final rules (seeds
    0 .. proxy-1), as s4_selection/ga.cu:109-198 computes them."""
    ec = evo or G.EvolutionConfig()
    vocab = vocab_of(genome, universe)
    tc = G.TreeConfig(vocab, list(FUNCTIONS))
    mask = term_mask_from(vocab)
    T = ec.train_samples
    seeds = list(range(proxy))
    dur = ctx.inst.durations
    gp = DeviceGP(ctx, tc, ec, seeds=seeds, T=T, term_mask=mask)
    try:
        gp.init()
        rows = np.concatenate([gen_durations(dur, T, s * 10000 + g) for g in range(pgens) for s in seeds])
        gp.run(rows, generations=pgens)
        fit = gp.evaluate(np.concatenate([gen_durations(dur, T, s * 10000 + pgens) for s in seeds]))
        tok, ln = gp.tokens()
        best = [int(np.argmax(fit[s])) for s in range(proxy)]              # first maximum (:177-181)
        finals = [tokens_to_tree(tok[s, b, :ln[s, b]], tc.terminals, tc.functions) for s, b in enumerate(best)]
    finally:
        gp.close()
    blk = gen_durations(dur, N_TRAIN, TRAIN_SEED)
    _, means = ctx.evaluate_means(pack_trees(finals), blk, N_TRAIN, mask)
    return means


def proxy_batch(ctx: FactoryContext, universe: Sequence[str], proxy: int = 16, pgens: int = 51):
    """A proxy function for sigp.ga_search.ga_search: genomes -> their
    per-seed values (the tool's batch_score, s4_selection/ga.cu:109-198, at
    `proxy` seeds and `pgens` generations), one genome's seeds side by side
    on the device at a time."""
    def run(genomes):
        return [list(proxy_scores(ctx, universe, g, proxy=proxy, pgens=pgens)) for g in genomes]
    return run
