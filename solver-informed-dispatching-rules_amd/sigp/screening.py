"""Surrogate pre-screening experiments (no counterpart in the reference).

    python -m sigp.screening --rules 100000 --generations 20

1. Labels: `--rules` random rules (ramped half-and-half over NS-11 +
   EXTRA-18) simulated exactly over 50 training scenarios; their
   phenotypes (64 situations x 8 candidates) computed on the device.
2. Network: the 512-256-256-1 MLP trained on 80 % of them; Spearman rank
   correlation and mean absolute error against the exact rewards on the
   other 20 %.
3. Screening: from a pool of fresh rules, the network's best eighth against
   a random eighth, both scored exactly -- what the network adds over
   picking without it.
4. GP: plain GP (each generation's 1,024 offspring simulated exactly) against
   surrogate-assisted GP (8,192 offspring bred, the network keeps the best
   1,024 for exact simulation), same instance, seeds and scenario schedule;
   held-out reward (500 scenarios) of each generation's best rule, and
   wall time split into host GP, exact simulation and the network.

Prints one JSON object.  Every reward is computed by the exact simulator.
"""
from __future__ import annotations

import argparse
import json
import random
import sys
import time
from typing import List

import numpy as np

from .api import FactoryContext
from .gp import EvolutionConfig, TreeConfig, gp_step, ramped_half_and_half
from .instance import Instance, synthetic_clustered
from .rules import EXTRA, FUNCTIONS, NS_TERMINALS, Node, pack_trees, term_mask_from
from .scenarios import gen_durations
from .surrogate import Mlp, features, make_situations, phenotype, spearman
from .workloads import TEST_SEED

VOCAB = NS_TERMINALS + EXTRA


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


class Scorer:
    """Exact mean rewards and surrogate scores of rule lists on one instance."""

    def __init__(self, inst: Instance, n_train: int = 50):
        self.inst = inst
        self.ctx = FactoryContext(inst)
        self.mask = term_mask_from(VOCAB)
        self.sit = make_situations(inst)
        self.n_train = n_train
        self.test = gen_durations(inst.durations, 500, TEST_SEED, mode="gamma")
        self.t_exact = 0.0
        self.t_net = 0.0

    def exact(self, trees: List[Node], scen: np.ndarray) -> np.ndarray:
        t0 = time.perf_counter()
        pk = pack_trees(trees)
        _, mean = self.ctx.evaluate_means(pk, scen, scen.shape[0], self.mask)
        self.t_exact += time.perf_counter() - t0
        return mean

    def train_scenarios(self, gen: int) -> np.ndarray:
        return gen_durations(self.inst.durations, self.n_train, 2_000_000 + 1000 * gen, mode="gamma")

    def features(self, trees: List[Node]) -> np.ndarray:
        t0 = time.perf_counter()
        X = features(phenotype(self.ctx, pack_trees(trees), self.mask, self.sit))
        self.t_net += time.perf_counter() - t0
        return X


def train_network(sc: Scorer, n_rules: int, epochs: int, seed: int):
    tc = TreeConfig(list(VOCAB), list(FUNCTIONS), 8, 2)
    trees = ramped_half_and_half(tc, random.Random(seed), n_rules)
    scen = sc.train_scenarios(0)
    t0 = time.perf_counter()
    y = sc.exact(trees, scen)
    t_label = time.perf_counter() - t0
    t0 = time.perf_counter()
    X = sc.features(trees)
    t_pheno = time.perf_counter() - t0
    mu, sd = float(y.mean()), float(y.std()) or 1.0
    ntr = int(0.8 * n_rules)
    net = Mlp(seed=seed)
    t0 = time.perf_counter()
    losses = net.train(X[:ntr], ((y[:ntr] - mu) / sd).astype(np.float32), epochs=epochs, batch=256,
                       lr=1e-3, seed=seed + 1)
    t_train = time.perf_counter() - t0
    pred = net.predict(X[ntr:]) * sd + mu
    stats = {"rules": n_rules, "train_rules": ntr, "epochs": epochs,
             "spearman": round(spearman(pred, y[ntr:]), 4),
             "mae": round(float(np.abs(pred - y[ntr:]).mean()), 3),
             "reward_sd": round(sd, 3), "final_loss": round(float(losses[-1]), 5),
             "seconds": {"labels": round(t_label, 3), "phenotypes": round(t_pheno, 3),
                         "training": round(t_train, 3)}}
    return net, (mu, sd), stats


def screening_gain(sc: Scorer, net: Mlp, norm, pool: int, seed: int):
    tc = TreeConfig(list(VOCAB), list(FUNCTIONS), 8, 2)
    trees = ramped_half_and_half(tc, random.Random(seed), pool)
    pred = net.predict(sc.features(trees)) * norm[1] + norm[0]
    true = sc.exact(trees, sc.train_scenarios(1))
    k = pool // 8
    top = np.argsort(-pred, kind="stable")[:k]
    rnd = np.random.default_rng(seed).choice(pool, k, replace=False)
    return {"pool": pool, "kept": k, "mean_kept_by_network": round(float(true[top].mean()), 3),
            "mean_kept_at_random": round(float(true[rnd].mean()), 3),
            "mean_best_possible": round(float(np.sort(true)[-k:].mean()), 3),
            "spearman_pool": round(spearman(pred, true), 4)}


def run_gp(sc: Scorer, net, norm, generations: int, pop: int, breed: int, seed: int, label: str):
    """One GP run; with net, each generation breeds `breed` offspring and the
    network keeps the best `pop` for exact simulation."""
    tc = TreeConfig(list(VOCAB), list(FUNCTIONS), 8, 2)
    ec = EvolutionConfig()
    rng = random.Random(seed)
    t_host = 0.0
    t0 = time.perf_counter()
    popl = ramped_half_and_half(tc, rng, pop)
    t_host += time.perf_counter() - t0
    curve = []
    wall0 = time.perf_counter()
    for gen in range(generations):
        fit = sc.exact(popl, sc.train_scenarios(gen)).tolist()
        best = popl[int(np.argmax(fit))]
        held = float(sc.exact([best], sc.test)[0])
        curve.append(round(held, 3))
        _log("%s gen %d best held-out %.2f" % (label, gen, held))
        if gen == generations - 1:
            break
        t0 = time.perf_counter()
        if net is None:
            popl = gp_step(popl, fit, rng, tc, ec)
            t_host += time.perf_counter() - t0
        else:
            pool: List[Node] = []
            while len(pool) < breed:
                pool.extend(gp_step(popl, fit, rng, tc, ec))
            t_host += time.perf_counter() - t0
            pred = net.predict(sc.features(pool)) * norm[1] + norm[0]
            keep = np.argsort(-pred, kind="stable")[:pop]
            popl = [pool[int(i)] for i in keep]
    return {"held_out_best_per_generation": curve, "wall_s": round(time.perf_counter() - wall0, 2),
            "host_gp_s": round(t_host, 2)}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--rules", type=int, default=100_000)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--pool", type=int, default=8192)
    ap.add_argument("--generations", type=int, default=20)
    ap.add_argument("--pop", type=int, default=1024)
    ap.add_argument("--breed", type=int, default=8192)
    ap.add_argument("--seed", type=int, default=1)
    a = ap.parse_args(argv)
    inst = synthetic_clustered(n=100, m=3, seed=1)
    sc = Scorer(inst)
    net, norm, stats = train_network(sc, a.rules, a.epochs, a.seed)
    _log("network", json.dumps(stats))
    gain = screening_gain(sc, net, norm, a.pool, a.seed + 100)
    _log("screening", json.dumps(gain))
    out = {"instance": "clustered n=100 m=3 (windows U[30,120])", "vocabulary": "NS-11 + EXTRA-18",
           "train_scenarios": 50, "held_out_scenarios": 500, "network": stats, "screening": gain}
    if a.generations > 0:
        sc.t_exact = sc.t_net = 0.0
        plain = run_gp(sc, None, norm, a.generations, a.pop, a.pop, a.seed + 7, "plain")
        plain.update(exact_s=round(sc.t_exact, 2))
        sc.t_exact = sc.t_net = 0.0
        surr = run_gp(sc, net, norm, a.generations, a.pop, a.breed, a.seed + 7, "surrogate")
        surr.update(exact_s=round(sc.t_exact, 2), network_s=round(sc.t_net, 2))
        target = plain["held_out_best_per_generation"][-1]
        reach = next((g for g, v in enumerate(surr["held_out_best_per_generation"]) if v >= 0.995 * target), None)
        out["gp"] = {"generations": a.generations, "pop": a.pop, "breed": a.breed, "plain": plain,
                     "surrogate": surr, "surrogate_generation_reaching_plain_final_within_0.5pct": reach}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
