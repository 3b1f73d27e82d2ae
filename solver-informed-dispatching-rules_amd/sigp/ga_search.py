"""The GA terminal-subset search (s4_selection/ga.cu:267-520) over a proxy
fitness function.

A genome is a bit set over the terminal universe; its proxy fitness is the
list of per-seed train-block means of GP trained on NS-11 plus the genome's
terminals (sigp.ga_device.proxy_scores computes it with every seed's GP loop
on the device).  The search is host logic on genomes, restated here step by
step so that, fed the same proxy values, it makes the same random draws
(one CPython-compatible Mersenne Twister seeded 20260714, ga.cu:87) and the
same decisions as the tool:

* pool (ga.cu:362-384): the neutral seeds -- the empty genome, the full
  universe unless nofull, every single terminal -- then random genomes (each
  bit with probability 0.35) up to 13; all evaluated, sorted by mean fitness
  (descending, ties by fewer terminals), cut to 13;
* a ternary tree of 13 (slot 0 the root, 1-3 the mids, 4-12 the leaves),
  whose clusters are restored by swapping a leader with its best supporter
  while one is better (ga.cu:386-407);
* each generation (ga.cu:431-478): when 6 or fewer genomes are distinct,
  two random genomes replace slots 12 and 11; every leader crosses with each
  of its three supporters (uniform crossover, then per-gene mutation with
  probability pm / L), the 12 children are evaluated, and a child replaces
  the worse of its two parents if it beats it; stop after `stag` generations
  without a better root;
* the distinct finalists, best first, are cut to `top`, re-scored at
  `confirm` seeds and `confirm_gens` generations, and the champion is the
  best confirmed mean (ties by fewer terminals) (ga.cu:487-511).

Evaluations go through an evaluation cache (genome -> per-seed values; the
"P" and "C" lines of the tool's --cache file, ga.cu:301-361), so a search
fed a stored cache replays the stored run.  The printed lines follow the
tool's format so that they compare with its logs.
"""
from __future__ import annotations

import random
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence

NP = 13
GA_SEED = 20260714          # s4_selection/ga.cu:87


def mean_of(v: Sequence[float]) -> float:
    """Sequential sum / count (ga.cu:295-299)."""
    a = 0.0
    for x in v:
        a += x
    return a / len(v)


def nbits(g: int) -> int:
    return bin(g).count("1")


def name_of(g: int, universe: Sequence[str]) -> str:
    s = ",".join(t for b, t in enumerate(universe) if g >> b & 1)
    return s if s else "(base)"


@dataclass
class SearchResult:
    champion: int
    champion_mean: float
    lines: List[str]
    evals: int
    finalists: List[int]
    cache: Dict[int, List[float]] = field(default_factory=dict)
    cache_conf: Dict[int, List[float]] = field(default_factory=dict)


def ga_search(universe: Sequence[str], proxy: Callable[[List[int]], List[List[float]]],
              confirm_fn: Callable[[List[int]], List[List[float]]], seed_mode: str = "random",
              nofull: bool = False, pm: float = 0.5, gens: int = 50, stag_k: int = 5, top_n: int = 4,
              cache: Optional[Dict[int, List[float]]] = None,
              cache_conf: Optional[Dict[int, List[float]]] = None,
              log: Optional[Callable[[str], None]] = None) -> SearchResult:
    """proxy(genomes) / confirm_fn(genomes) return the per-seed values of
    each genome (the tool's batch_score at proxy and at confirm seeds);
    they are called only for genomes missing from the caches."""
    nb = len(universe)
    cache = {} if cache is None else dict(cache)
    cache_conf = {} if cache_conf is None else dict(cache_conf)
    rng = random.Random(GA_SEED)
    lines: List[str] = []
    evals = 0

    def emit(s: str) -> None:
        lines.append(s)
        if log:
            log(s)

    def m(g: int) -> float:
        return mean_of(cache[g])

    def ensure(want: List[int]) -> None:          # ga.cu:347-360
        nonlocal evals
        miss = [g for g in sorted(set(want)) if g not in cache]
        if not miss:
            return
        res = proxy(miss)
        for g, v in zip(miss, res):
            cache[g] = list(v)
        evals += len(miss)
        emit("  batch %d genomes" % len(miss))

    def rand_genome() -> int:
        g = 0
        for b in range(nb):
            if rng.random() < 0.35:
                g |= 1 << b
        return g

    full = (1 << nb) - 1
    seeds: List[int] = []
    if seed_mode == "neutral":
        seeds.append(0)
        if not nofull:
            seeds.append(full)
        seeds += [1 << b for b in range(nb)]
    seeds = sorted(set(seeds))
    pool = list(seeds)
    while len(pool) < NP:
        pool.append(rand_genome())
    ensure(pool)
    # stable sort by mean (descending), then fewer terminals
    pool = sorted(pool, key=lambda g: (-m(g), nbits(g)))[:NP]
    tree = list(pool)

    def restore() -> None:                        # ga.cu:386-405
        for _ in range(NP):
            moved = False
            for c in (3, 2, 1, 0):
                lead = 0 if c == 0 else c
                base = 1 if c == 0 else 4 + 3 * (c - 1)
                best = -1
                for j in range(3):
                    idx = base + j
                    if m(tree[idx]) > m(tree[lead]):
                        if best < 0 or m(tree[idx]) > m(tree[best]):
                            best = idx
                if best >= 0:
                    tree[lead], tree[best] = tree[best], tree[lead]
                    moved = True
            if not moved:
                break

    restore()

    def cross_mut(a: int, b: int) -> int:         # ga.cu:420-427
        child = 0
        for k in range(nb):
            child |= (a if rng.random() < 0.5 else b) & (1 << k)
        for k in range(nb):
            if rng.random() < pm / nb:
                child ^= 1 << k
        return child

    best_seen = m(tree[0])
    stag = 0
    for gen in range(gens):
        pmean = 0.0
        for g in tree:
            pmean += m(g)
        pmean /= NP
        uniq = len(set(tree))
        emit("gen %d | best %.2f [%s] | mean %.2f | unique %d | evals %d"
             % (gen, m(tree[0]), name_of(tree[0], universe), pmean, uniq, evals))
        if uniq <= 6:
            fresh = []
            for r in range(2):
                g = rand_genome()
                tree[12 - r] = g
                fresh.append(g)
            ensure(fresh)
            emit("  reinject 2")
        children, pairs = [], []
        for c in range(4):
            lead = 0 if c == 0 else c
            base = 1 if c == 0 else 4 + 3 * (c - 1)
            for j in range(3):
                sup = base + j
                children.append(cross_mut(tree[lead], tree[sup]))
                pairs.append((lead, sup))
        ensure(children)
        for ch, (lead, sup) in zip(children, pairs):
            wp = sup if m(tree[sup]) <= m(tree[lead]) else lead
            if m(ch) > m(tree[wp]):
                tree[wp] = ch
        restore()
        cur = m(tree[0])
        if cur > best_seen:
            best_seen, stag = cur, 0
        else:
            stag += 1
        if stag >= stag_k:
            emit("stop: stagnant %d gens" % stag)
            break
    pmean = 0.0
    for g in tree:
        pmean += m(g)
    pmean /= NP
    emit("gen end | best %.2f [%s] | mean %.2f | evals %d" % (m(tree[0]), name_of(tree[0], universe), pmean, evals))

    # distinct finalists in genome order, then stable by mean and size (ga.cu:487-494)
    fin = sorted(sorted(set(tree)), key=lambda g: (-m(g), nbits(g)))[:top_n]
    cmiss = [g for g in fin if g not in cache_conf]
    if cmiss:
        for g, v in zip(cmiss, confirm_fn(cmiss)):
            cache_conf[g] = list(v)
        emit("  batch %d genomes (confirm)" % len(cmiss))
    bi = 0
    for i in range(1, len(fin)):
        a, b = mean_of(cache_conf[fin[i]]), mean_of(cache_conf[fin[bi]])
        if a > b or (a == b and nbits(fin[i]) < nbits(fin[bi])):
            bi = i
    for g in fin:
        emit("  %.2f (proxy %.2f) [%s]" % (mean_of(cache_conf[g]), m(g), name_of(g, universe)))
    emit("CHAMPION train-%d %.2f | evals %d | terms: %s" % (len(cache_conf[fin[bi]]), mean_of(cache_conf[fin[bi]]),
                                                            evals + len(cmiss), name_of(fin[bi], universe)))
    return SearchResult(fin[bi], mean_of(cache_conf[fin[bi]]), lines, evals + len(cmiss), fin, cache, cache_conf)


def read_cache(lines) -> tuple:
    """The tool's cache file (ga.cu:301-361): header, then "P <hex> <n>
    <hexfloat>..." (proxy) and "C ..." (confirm) lines, in append order.
    Returns (header, [(kind, genome, values)])."""
    it = iter(lines)
    header = next(it).rstrip("\n")
    entries = []
    for ln in it:
        p = ln.split()
        if len(p) < 3 or p[0] not in ("P", "C"):
            continue
        n = int(p[2])
        vals = [float.fromhex(x) for x in p[3:3 + n]]
        if len(vals) == n:
            entries.append((p[0], int(p[1], 16), vals))
    return header, entries
