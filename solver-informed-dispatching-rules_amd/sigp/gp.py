"""GP operators of the SI-GP loop, driven by CPython's Mersenne Twister.

Restates engine/gp.hpp:13-194: ramped half-and-half initialisation, Koza
90/10 crossover-point selection, subtree crossover and mutation under the
depth limit, size-7 tournament selection and the generation step with
elitism.  engine/rng.hpp restates CPython's `random.Random` (init_by_array
seeding, 53-bit random(), _randbelow by getrandbits rejection, sample's
set-size rule), so a `random.Random(seed)` consumed in the same order gives
the same populations; tests pin this against verify/golden/gp.txt.
"""
from __future__ import annotations

import random
from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

from .rules import FUNCTIONS, Node, tree_depth

__all__ = ["TreeConfig", "EvolutionConfig", "ramped_half_and_half", "gp_step",
           "tournament_selection", "subtree_crossover", "subtree_mutation", "generate_full",
           "generate_grow", "select_crossover_point", "get_subtree", "replace_subtree"]


@dataclass
class TreeConfig:
    """engine/gp.hpp:13-18."""

    terminals: List[str]
    functions: List[str] = field(default_factory=lambda: list(FUNCTIONS))
    max_depth: int = 8
    min_depth: int = 2


@dataclass
class EvolutionConfig:
    """engine/gp.hpp:20-34."""

    population_size: int = 1024
    generations: int = 51
    max_depth: int = 8
    min_depth: int = 2
    tournament_size: int = 7
    elitism: int = 10
    crossover_rate: float = 0.80
    mutation_rate: float = 0.15
    reproduction_rate: float = 0.05
    num_vehicles: int = 3
    sigma: float = 0.2
    train_samples: int = 1
    test_samples: int = 500


def _below(rng: random.Random, n: int) -> int:
    """engine/rng.hpp:87-93 (randbelow; 0 for n == 0)."""
    return 0 if n == 0 else rng._randbelow(n)


def pick_terminal(cfg: TreeConfig, rng: random.Random) -> Node:
    """engine/gp.hpp:37-45: 'ERC' draws a constant printed with %.3f."""
    t = cfg.terminals[_below(rng, len(cfg.terminals))]
    if t == "ERC":
        return Node("%.3f" % rng.random())
    return Node(t)


def generate_full(cfg: TreeConfig, rng: random.Random, depth: int) -> Node:
    """engine/gp.hpp:47-53."""
    if depth <= 0:
        return pick_terminal(cfg, rng)
    sym = cfg.functions[_below(rng, len(cfg.functions))]
    a = generate_full(cfg, rng, depth - 1)
    b = generate_full(cfg, rng, depth - 1)
    return Node(sym, (a, b))


def generate_grow(cfg: TreeConfig, rng: random.Random, depth: int) -> Node:
    """engine/gp.hpp:55-64."""
    if depth <= 0:
        return pick_terminal(cfg, rng)
    nt, nf = float(len(cfg.terminals)), float(len(cfg.functions))
    if rng.random() < nt / (nt + nf):
        return pick_terminal(cfg, rng)
    sym = cfg.functions[_below(rng, len(cfg.functions))]
    a = generate_grow(cfg, rng, depth - 1)
    b = generate_grow(cfg, rng, depth - 1)
    return Node(sym, (a, b))


def ramped_half_and_half(cfg: TreeConfig, rng: random.Random, pop_size: int) -> List[Node]:
    """engine/gp.hpp:67-77: depths ramp min..max, alternating full/grow blocks."""
    nd = cfg.max_depth - cfg.min_depth + 1
    pop = []
    for i in range(pop_size):
        d = cfg.min_depth + i % nd
        pop.append(generate_full(cfg, rng, d) if (i // nd) % 2 == 0 else generate_grow(cfg, rng, d))
    return pop


def _preorder(root: Node) -> List[Node]:
    out: List[Node] = []
    stack = [root]
    while stack:
        n = stack.pop()
        out.append(n)
        stack.extend(reversed(n.children))
    return out


def get_subtree(root: Node, index: int) -> Node:
    """engine/gp.hpp:90-102: node at pre-order position `index`."""
    return _preorder(root)[index]


def replace_subtree(root: Node, index: int, repl: Node) -> Node:
    """engine/gp.hpp:104-116."""
    counter = [-1]

    def rec(n: Node) -> Node:
        counter[0] += 1
        if counter[0] == index:
            return repl
        if n.is_terminal:
            return n
        return Node(n.label, tuple(rec(c) for c in n.children))

    return rec(root)


def select_crossover_point(tree: Node, rng: random.Random, p_internal: float = 0.9) -> int:
    """engine/gp.hpp:120-131: Koza 90/10; the rng is consumed only when both
    node classes exist."""
    nodes = _preorder(tree)
    internals = [i for i, n in enumerate(nodes) if not n.is_terminal]
    leaves = [i for i, n in enumerate(nodes) if n.is_terminal]
    if internals and (not leaves or rng.random() < p_internal):
        return internals[_below(rng, len(internals))]
    if leaves:
        return leaves[_below(rng, len(leaves))]
    return internals[_below(rng, len(internals))]


def subtree_crossover(p1: Node, p2: Node, rng: random.Random, max_depth: int) -> Tuple[Node, Node]:
    """engine/gp.hpp:133-144."""
    pt1 = select_crossover_point(p1, rng)
    pt2 = select_crossover_point(p2, rng)
    s1 = get_subtree(p1, pt1)
    s2 = get_subtree(p2, pt2)
    c1 = replace_subtree(p1, pt1, s2)
    c2 = replace_subtree(p2, pt2, s1)
    if tree_depth(c1) > max_depth:
        c1 = p1
    if tree_depth(c2) > max_depth:
        c2 = p2
    return c1, c2


def subtree_mutation(ind: Node, cfg: TreeConfig, rng: random.Random) -> Node:
    """engine/gp.hpp:146-153."""
    pt = select_crossover_point(ind, rng)
    nd = cfg.min_depth + _below(rng, cfg.max_depth - cfg.min_depth + 1)   # randint
    sub = generate_grow(cfg, rng, nd)
    mut = replace_subtree(ind, pt, sub)
    return ind if tree_depth(mut) > cfg.max_depth else mut


def tournament_selection(pop: Sequence[Node], fit: Sequence[float], size: int,
                         rng: random.Random) -> Node:
    """engine/gp.hpp:157-166: k distinct contenders, first maximum wins."""
    k = min(size, len(pop))
    cont = rng.sample(range(len(pop)), k)
    best = cont[0]
    for c in cont[1:]:
        if fit[c] > fit[best]:
            best = c
    return pop[best]


def gp_step(pop: Sequence[Node], fit: Sequence[float], rng: random.Random, tcfg: TreeConfig,
            cfg: EvolutionConfig) -> List[Node]:
    """engine/gp.hpp:170-194: stable descending sort, elitism, then
    crossover / mutation / reproduction by rate."""
    order = sorted(range(len(pop)), key=lambda i: fit[i], reverse=True)
    new = [pop[order[e]] for e in range(min(cfg.elitism, len(order)))]
    while len(new) < cfg.population_size:
        r = rng.random()
        if r < cfg.crossover_rate:
            a = tournament_selection(pop, fit, cfg.tournament_size, rng)
            b = tournament_selection(pop, fit, cfg.tournament_size, rng)
            c1, c2 = subtree_crossover(a, b, rng, cfg.max_depth)
            new.append(c1)
            if len(new) < cfg.population_size:
                new.append(c2)
        elif r < cfg.crossover_rate + cfg.mutation_rate:
            new.append(subtree_mutation(tournament_selection(pop, fit, cfg.tournament_size, rng),
                                        tcfg, rng))
        else:
            new.append(tournament_selection(pop, fit, cfg.tournament_size, rng))
    return new[:cfg.population_size]
