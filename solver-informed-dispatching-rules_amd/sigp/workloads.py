"""Named workloads: the synthetic instance, rule set and scenario set of each
benchmark configuration, as pure functions of fixed seeds
(instance seed 1, rules seed 0, scenarios base 1,000,000 + s).
"""
from __future__ import annotations

import hashlib
import random
from dataclasses import dataclass
from typing import List

import numpy as np

from .gp import TreeConfig, ramped_half_and_half
from .instance import Instance, synthetic_clustered, synthetic_mixed, synthetic_uniform
from .rules import EXTRA, FUNCTIONS, NS_TERMINALS, Node, Packed, pack_trees, term_mask_from
from .scenarios import gen_durations

TEST_SEED = 1_000_000   # engine/gp_engine.cu:22

__all__ = ["Workload", "MultiWorkload", "plain_small", "large", "gp_generation",
           "gp_reference_generation", "batched_uneven", "random_rules", "workload_digest", "TEST_SEED"]


@dataclass
class Workload:
    name: str
    inst: Instance
    trees: List[Node]
    packed: Packed
    mask: np.ndarray
    sampled: np.ndarray      # rows of n durations
    n_scen: int              # scenarios per rule (T)
    group_size: int = 0      # rule p reads rows (p // group_size) * T + s; 0 = one group

    @property
    def n_sims(self) -> int:
        return self.packed.n_rules * self.n_scen


def random_rules(n_rules: int, terminals, seed: int = 0, min_depth: int = 1,
                 max_depth: int = 5, functions=FUNCTIONS) -> List[Node]:
    """Ramped half-and-half rules (engine/gp.hpp:67-77); depths in the
    reference convention (a leaf has depth 0), so 1..5 = 2..6 levels, <= 63 nodes."""
    cfg = TreeConfig(list(terminals), list(functions), max_depth, min_depth)
    return ramped_half_and_half(cfg, random.Random(seed), n_rules)


def plain_small(n_rules: int = 1024, n_scen: int = 500, n: int = 100, m: int = 2,
                mode: str = "gamma", shape: float = 4.0) -> Workload:
    """1 uniform instance (n=100, m=2), EXTRA-18 rules, Gamma(4, mean 10) scenarios."""
    inst = synthetic_uniform(n=n, m=m, seed=1)
    trees = random_rules(n_rules, EXTRA, seed=0)
    sampled = gen_durations(inst.durations, n_scen, TEST_SEED, mode=mode, shape=shape)
    return Workload("plain-small n=%d m=%d %dx%d" % (n, m, n_rules, n_scen), inst, trees,
                    pack_trees(trees), term_mask_from(EXTRA), sampled, n_scen)


def large(n_rules: int = 4096, n_scen: int = 2000, n: int = 288, m: int = 4) -> Workload:
    """288-customer clustered+uniform instance, m=4, rules <= 127 nodes, Gamma(2)."""
    inst = synthetic_mixed(n=n, m=m, seed=1)
    trees = random_rules(n_rules, EXTRA, seed=0, min_depth=1, max_depth=6)
    sampled = gen_durations(inst.durations, n_scen, TEST_SEED, mode="gamma", shape=2.0)
    return Workload("large n=%d m=%d %dx%d" % (n, m, n_rules, n_scen), inst, trees,
                    pack_trees(trees), term_mask_from(EXTRA), sampled, n_scen)


def gp_generation(pop: int = 1024, n_scen: int = 50) -> Workload:
    """One generation of the on-device GP loop: the clustered instance
    (n=100, m=3, windows U[30,120]), a ramped half-and-half population over
    NS-11 + EXTRA-18 (max depth 8), 50 common training scenarios (Gamma 4)."""
    inst = clustered_gp_instance()
    trees = random_rules(pop, NS_TERMINALS + EXTRA, seed=0, min_depth=2, max_depth=8)
    sampled = gen_durations(inst.durations, n_scen, TEST_SEED, mode="gamma", shape=4.0)
    return Workload("gp-generation n=100 m=3 %dx%d" % (pop, n_scen), inst, trees,
                    pack_trees(trees), ext29_mask(), sampled, n_scen)


def gp_reference_generation(seeds: int = 30, pop: int = 1024) -> Workload:
    """One launch of the reference's GP loop (engine/gp_engine.cu:155-164):
    `seeds` independent populations, one training scenario each (Mei
    protocol, engine/gp.hpp:32), rule p reading scenario row p // pop.
    n_scen is 1; the scenario rows are one per seed."""
    inst = clustered_gp_instance()
    trees = random_rules(seeds * pop, NS_TERMINALS + EXTRA, seed=0, min_depth=2, max_depth=8)
    sampled = np.concatenate([gen_durations(inst.durations, 1, s * 10000, mode="gamma", shape=4.0)
                              for s in range(seeds)])
    return Workload("gp-reference-generation n=100 m=3 %d seeds x %d rules x 1" % (seeds, pop),
                    inst, trees, pack_trees(trees), ext29_mask(), sampled, 1, pop)


@dataclass
class MultiWorkload:
    name: str
    insts: List[Instance]
    trees: List[Node]
    packed: Packed
    mask: np.ndarray
    blocks: List[np.ndarray]   # per instance: T x n_k scenario rows
    n_scen: int

    @property
    def n_sims(self) -> int:
        return len(self.insts) * self.packed.n_rules * self.n_scen


def batched_uneven(n_inst: int = 40, n_rules: int = 256, n_scen: int = 500, seed: int = 1) -> MultiWorkload:
    """Many uneven instances in one batch: n drawn from {48, 96, ..., 288},
    uniform and clustered coordinates alternating, m in {1..4}, windows
    U[10,200]; ramped half-and-half rules of 3-127 nodes over EXTRA-18; each
    instance with its own Gamma(4, mean 10) scenarios (base 1,000,000)."""
    rng = random.Random(seed)
    insts = []
    for k in range(n_inst):
        n = rng.choice([48, 96, 144, 192, 240, 288])
        m = rng.randint(1, 4)
        gen = synthetic_uniform if k % 2 == 0 else synthetic_clustered
        insts.append(gen(n=n, m=m, seed=100 + k, width=(10.0, 200.0)))
    trees = random_rules(n_rules, EXTRA, seed=0, min_depth=1, max_depth=6)
    blocks = [gen_durations(inst.durations, n_scen, TEST_SEED, mode="gamma", shape=4.0) for inst in insts]
    return MultiWorkload("batched-uneven %d instances x %d rules x %d" % (n_inst, n_rules, n_scen),
                         insts, trees, pack_trees(trees), term_mask_from(EXTRA), blocks, n_scen)


def ext29_mask() -> np.ndarray:
    return term_mask_from(NS_TERMINALS + EXTRA)


def clustered_gp_instance() -> Instance:
    return synthetic_clustered(n=100, m=3, seed=1)


def workload_digest(wl) -> str:
    """sha256 over a workload's arrays: instance(s), packed rules, mask and
    scenario rows.  A committed work model (operations per simulation) is
    used for a workload only when its digest matches."""
    h = hashlib.sha256()
    insts = wl.insts if isinstance(wl, MultiWorkload) else [wl.inst]
    for inst in insts:
        for a in (inst.scores, inst.durations, inst.open_t, inst.close_t, inst.dist):
            h.update(np.ascontiguousarray(a).tobytes())
    for a in (wl.packed.ops, wl.packed.args, wl.packed.offsets, wl.packed.lengths, wl.mask):
        h.update(np.ascontiguousarray(a).tobytes())
    for rows in (wl.blocks if isinstance(wl, MultiWorkload) else [wl.sampled]):
        h.update(np.ascontiguousarray(rows).tobytes())
    return h.hexdigest()
