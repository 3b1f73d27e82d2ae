"""Surrogate pre-screening experiments (no counterpart in the reference).

    python -m sigp.screening --rules 100000 --generations 50

1. Labels: `--rules` random rules (ramped half-and-half over NS-11 +
   EXTRA-18) simulated exactly over 50 label scenarios; their phenotypes
   (64 situations x 8 candidates) computed on the device.
2. Network: the 512-256-256-1 MLP trained on 80 % of them; Spearman rank
   correlation and mean absolute error against the exact rewards on the
   other 20 %.
3. Screening: from a pool of fresh rules, the network's best eighth against
   a random eighth, both scored exactly -- what the network adds over
   picking without it.
4. GP, both loops entirely on the device: plain GP (each generation's 1,024
   offspring simulated exactly on 50 training scenarios) against
   surrogate-assisted GP (8 x 1,024 offspring bred, the network keeps the
   best 1,024 for exact simulation), same instance, seeds and scenario
   schedule.  Each generation's best rule (by its training fitness) is
   scored on 500 held-out scenarios; the result is the device time at which
   each loop first reaches plain GP's final held-out reward within 0.5 %,
   compared with plain GP's whole run and with plain GP's own time to that
   band, each with and without the network's setup (labelling, phenotypes,
   training), which the runs sharing the network split between them.
   With --online there is no setup: each surrogate run starts with plain
   steps and trains its own network from the phenotypes and exact fitness
   of the populations it evaluates anyway (run_device_gp_online), and that
   time is part of the run's.

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
from .gp import EvolutionConfig, TreeConfig, ramped_half_and_half
from .instance import Instance, synthetic_clustered, synthetic_mixed
from .rules import EXTRA, FUNCTIONS, NS_TERMINALS, Node, pack_trees, term_mask_from
from .scenarios import gen_durations
from .surrogate import Mlp, features, make_situations, phenotype, spearman
from .workloads import TEST_SEED

VOCAB = NS_TERMINALS + EXTRA


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


class Scorer:
    """Exact mean rewards and surrogate scores of rule lists on one instance."""

    def __init__(self, inst: Instance, n_train: int = 50, shape: float = 4.0, depth: int = 8):
        self.inst = inst
        self.ctx = FactoryContext(inst)
        self.mask = term_mask_from(VOCAB)
        self.sit = make_situations(inst)
        self.n_train = n_train
        self.shape = shape
        self.depth = depth          # TreeConfig / EvolutionConfig max_depth
        self.test = gen_durations(inst.durations, 500, TEST_SEED, mode="gamma", shape=shape)
        self.t_exact = 0.0
        self.t_net = 0.0

    def exact(self, trees: List[Node], scen: np.ndarray) -> np.ndarray:
        t0 = time.perf_counter()
        pk = pack_trees(trees)
        _, mean = self.ctx.evaluate_means(pk, scen, scen.shape[0], self.mask)
        self.t_exact += time.perf_counter() - t0
        return mean

    def train_scenarios(self, gen: int) -> np.ndarray:
        return gen_durations(self.inst.durations, self.n_train, 2_000_000 + 1000 * gen, mode="gamma",
                             shape=self.shape)

    def features(self, trees: List[Node]) -> np.ndarray:
        t0 = time.perf_counter()
        X = features(phenotype(self.ctx, pack_trees(trees), self.mask, self.sit))
        self.t_net += time.perf_counter() - t0
        return X


def device_label_set(sc: Scorer, n_rules: int, scen: np.ndarray, seed: int, pop_max: int = 4096):
    """A labelled set built on the device: ceil(n_rules / pop_max) GP
    populations initialised by ramped half-and-half (run r seeded
    seed * 1,000,003 + r), each tree's exact mean over `scen` and its
    phenotype, computed from the device's own compiled trees.  Returns
    (features, labels, label seconds, phenotype seconds)."""
    from .gp_device import DeviceGP
    runs = -(-n_rules // pop_max)
    pop = -(-n_rules // runs)
    t0 = time.perf_counter()
    tc = TreeConfig(list(VOCAB), list(FUNCTIONS), sc.depth, 2)
    ec = EvolutionConfig(population_size=pop, max_depth=sc.depth)
    gp = DeviceGP(sc.ctx, tc, ec, seeds=[seed * 1_000_003 + r for r in range(runs)], T=scen.shape[0],
                  term_mask=sc.mask)
    gp.init()
    y = gp.evaluate(np.tile(scen, (runs, 1))).reshape(-1)
    t_label = time.perf_counter() - t0
    t0 = time.perf_counter()
    X = features(gp.phenotype(sc.sit))
    t_pheno = time.perf_counter() - t0
    gp.close()
    return X, y, t_label, t_pheno


def train_network(sc: Scorer, n_rules: int, epochs: int, seed: int, label_scenarios: int = 0,
                  on_device: bool = True):
    """The network trained on random rules (ramped half-and-half) labelled by
    their exact mean over the first `label_scenarios` (0: all 50) label
    scenarios; held out: the last 20 % of them.  On the device the rules are
    generated, compiled, labelled and phenotyped there (device_label_set,
    n_rules rounded up to whole populations); on the host they are
    generated and compiled in Python.  Either way the label time counts the
    rules' generation."""
    scen = sc.train_scenarios(0)
    if label_scenarios:
        scen = np.ascontiguousarray(scen[:label_scenarios])
    if on_device:
        X, y, t_label, t_pheno = device_label_set(sc, n_rules, scen, seed)
        n_rules = y.shape[0]
    else:
        t0 = time.perf_counter()
        tc = TreeConfig(list(VOCAB), list(FUNCTIONS), sc.depth, 2)
        trees = ramped_half_and_half(tc, random.Random(seed), n_rules)
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
    stats = {"rules": n_rules, "train_rules": ntr, "epochs": epochs, "label_scenarios": int(scen.shape[0]),
             "labelled_on": "device" if on_device else "host",
             "spearman": round(spearman(pred, y[ntr:]), 4),
             "mae": round(float(np.abs(pred - y[ntr:]).mean()), 3),
             "reward_sd": round(sd, 3), "final_loss": round(float(losses[-1]), 5),
             "seconds": {"labels": round(t_label, 3), "phenotypes": round(t_pheno, 3),
                         "training": round(t_train, 3)}}
    return net, (mu, sd), stats


def screening_gain(sc: Scorer, net: Mlp, norm, pool: int, seed: int):
    tc = TreeConfig(list(VOCAB), list(FUNCTIONS), sc.depth, 2)
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


def run_device_gp(sc: Scorer, net, generations: int, pop: int, breed: int, seed: int, label: str,
                  chunk: int = 5, launch=None, screened_init: bool = False):
    """One GP loop on the device (one HIP graph per generation): plain, or
    screened by `net` when given (breed = k x pop offspring per generation),
    run in calls of `chunk` generations.  Returns the held-out reward of each
    generation's best rule and the device time at the end of each
    generation."""
    from .gp_device import DeviceGP
    tc = TreeConfig(list(VOCAB), list(FUNCTIONS), sc.depth, 2)
    ec = EvolutionConfig(population_size=pop, max_depth=sc.depth)
    T = sc.n_train
    gp = DeviceGP(sc.ctx, tc, ec, seeds=[seed], T=T, term_mask=sc.mask,
                  offspring=breed if net is not None else 0, **(launch or {}))
    if net is not None:
        gp.set_surrogate(net, sc.sit)
    t0 = time.perf_counter()
    gp.init(screened=screened_init and net is not None)
    init_ms = 1000.0 * (time.perf_counter() - t0) if screened_init and net is not None else 0.0
    best, gms, t_dev = [], [], init_ms
    for g0 in range(0, generations, chunk):
        k = min(chunk, generations - g0)
        rows = np.concatenate([sc.train_scenarios(g) for g in range(g0, g0 + k)])
        res = gp.run(rows, generations=k, graph=True, step_last=g0 + k < generations)
        best += [res.best[g][0] for g in range(k)]
        gms += [t_dev + float(v) for v in res.generation_ms]
        t_dev = gms[-1]
        if generations > chunk:
            _log("  %s: generation %d done, %.0f ms device time" % (label, g0 + k - 1, t_dev))
    held = sc.exact(best, sc.test)
    _log("%s: %d generations in %.1f ms; held-out of the last best %.2f"
         % (label, generations, t_dev, float(held[-1])))
    gp.close()
    return {"held_out_best_per_generation": [round(float(v), 3) for v in held],
            "generation_ms": [round(float(v), 1) for v in gms]}


def run_device_gp_online(sc: Scorer, generations: int, pop: int, breed: int, seed: int, label: str,
                         train_at=(2, 4, 8, 16), epochs: int = 20, cap: int = 16384, chunk: int = 5,
                         launch=None):
    """Surrogate-assisted GP whose network learns inside the run, from the
    run's own exact evaluations: until generation train_at[0] the loop takes
    plain steps; before each generation up to train_at[-1] the population's
    phenotypes are taken, and after its evaluation they are kept with their
    exact fitness; at each generation of `train_at` the network (a fresh one
    at the first) trains on the latest `cap` of them, and from the first on it
    screens breed = k x pop offspring per generation.  No rule is simulated
    for the network alone.  The time at the end of each generation is the
    device time plus the host time of the phenotypes and the training, both
    spent inside the run.  Returns what run_device_gp returns, plus the
    training record."""
    from .gp_device import DeviceGP
    tc = TreeConfig(list(VOCAB), list(FUNCTIONS), sc.depth, 2)
    ec = EvolutionConfig(population_size=pop, max_depth=sc.depth)
    gp = DeviceGP(sc.ctx, tc, ec, seeds=[seed], T=sc.n_train, term_mask=sc.mask, offspring=breed, **(launch or {}))
    gp.init()
    net = Mlp(seed=seed)
    X, Y = [], []
    best, gms, trained = [], [], []
    t_dev, t_host = 0.0, 0.0
    last = max(train_at)
    g = 0
    while g < generations:
        if g < last:
            t0 = time.perf_counter()
            Xg = features(phenotype(sc.ctx, gp.packed(), sc.mask, sc.sit))
            t_host += 1000.0 * (time.perf_counter() - t0)
            res = gp.run(sc.train_scenarios(g), generations=1, graph=False, step_last=g + 1 < generations)
            X.append(Xg)
            Y.append(res.fitness[0, 0].astype(np.float32))
            k = 1
        else:
            k = min(chunk, generations - g)
            rows = np.concatenate([sc.train_scenarios(q) for q in range(g, g + k)])
            res = gp.run(rows, generations=k, graph=True, step_last=g + k < generations)
        best += [res.best[q][0] for q in range(k)]
        base = t_dev
        for q in range(k):
            gms.append(base + float(res.generation_ms[q]) + t_host)
        t_dev = base + float(res.generation_ms[-1])
        g += k
        if g in train_at and g < generations:
            t0 = time.perf_counter()
            Xa = np.concatenate(X)[-cap:]
            Ya = np.concatenate(Y)[-cap:]
            mu, sd = float(Ya.mean()), float(Ya.std()) or 1.0
            losses = net.train(Xa, ((Ya - mu) / sd).astype(np.float32), epochs=epochs, batch=256, lr=1e-3,
                               seed=seed + g)
            if not trained:
                gp.set_surrogate(net, sc.sit)
            dt = 1000.0 * (time.perf_counter() - t0)
            t_host += dt
            trained.append({"generation": g, "rows": int(Xa.shape[0]), "ms": round(dt, 1),
                            "final_loss": round(float(losses[-1]), 5)})
            _log("  %s: network trained at generation %d on %d rules in %.0f ms" % (label, g, Xa.shape[0], dt))
    held = sc.exact(best, sc.test)
    _log("%s: %d generations in %.1f ms (host %.0f ms of it); held-out of the last best %.2f"
         % (label, generations, gms[-1], t_host, float(held[-1])))
    gp.close()
    net.close()
    return {"held_out_best_per_generation": [round(float(v), 3) for v in held],
            "generation_ms": [round(float(v), 1) for v in gms], "host_ms": round(t_host, 1), "training": trained}


def reach(curve, ms, target):
    """First generation whose held-out reward is within 0.5 % of `target`,
    and the device time at its end."""
    for g, v in enumerate(curve):
        if v >= 0.995 * target:
            return g, ms[g]
    return None, None


def ratios_of(sms, pms, total, setup_share_ms):
    """Surrogate-assisted GP's device time to plain GP's final held-out band
    (sms) against: plain GP's whole run (total), the same with the
    network's setup (labels, phenotypes, training; shared by the runs that
    use the network) added to the surrogate's time, and plain GP's own time
    to that band (pms), with and without the setup."""
    if sms is None:
        return {"time_ratio": None, "time_ratio_with_setup": None, "time_ratio_to_plain_reach": None,
                "time_ratio_to_plain_reach_with_setup": None}
    return {"time_ratio": round(sms / total, 4),
            "time_ratio_with_setup": round((sms + setup_share_ms) / total, 4),
            "time_ratio_to_plain_reach": round(sms / pms, 4),
            "time_ratio_to_plain_reach_with_setup": round((sms + setup_share_ms) / pms, 4)}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--rules", type=int, default=100_000)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--label-scenarios", type=int, default=0,
                    help="scenarios per labelled rule (0: the 50 training scenarios)")
    ap.add_argument("--labels-on-host", action="store_true",
                    help="generate and compile the labelled rules in Python instead of on the device")
    ap.add_argument("--pool", type=int, default=8192)
    ap.add_argument("--generations", type=int, default=50)
    ap.add_argument("--config", choices=["clustered", "large"], default="clustered")
    ap.add_argument("--waves", type=int, default=0, help="evaluation wavefronts per workgroup (0: automatic)")
    ap.add_argument("--mapping", choices=["scenarios", "candidates", "candidates2"], default="candidates")
    ap.add_argument("--pop", type=int, default=0, help="population (0: 1,024; large: 4,096)")
    ap.add_argument("--breed", type=int, default=0, help="offspring bred per generation (0: 8 x population)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--gp-runs", type=int, default=5, help="GP seed pairs (plain, surrogate) with one network")
    ap.add_argument("--screened-init", action="store_true",
                    help="surrogate runs start from a screened initial population (k ramped half-and-half "
                         "populations, the network keeps one; its time counts in the run)")
    ap.add_argument("--online", action="store_true",
                    help="train each surrogate run's network inside the run from its own exact evaluations "
                         "(no labelling setup) instead of one network trained beforehand")
    a = ap.parse_args(argv)
    if a.config == "large":
        # n=288 clustered + uniform, m=4, windows U[20,200], Gamma(2), rules <= 127 nodes
        inst = synthetic_mixed(n=288, m=4, seed=1)
        sc = Scorer(inst, shape=2.0, depth=6)
        desc = "clustered + uniform n=288 m=4 (windows U[20,200]), Gamma(2)"
        a.pop = a.pop or 4096
    else:
        inst = synthetic_clustered(n=100, m=3, seed=1)
        sc = Scorer(inst)
        desc = "clustered n=100 m=3 (windows U[30,120]), Gamma(4)"
        a.pop = a.pop or 1024
    a.breed = a.breed or 8 * a.pop
    launch = None
    if a.waves:
        launch = dict(waves_per_item=a.waves, mapping={"scenarios": 1, "candidates": 2,
                                                       "candidates2": 3}[a.mapping])
    out = {"instance": desc, "vocabulary": "NS-11 + EXTRA-18", "max_depth": sc.depth,
           "train_scenarios": 50, "held_out_scenarios": 500}
    if a.online:
        net, setup_s = None, 0.0
        out["network"] = "trained inside each surrogate run from its own exact evaluations"
    else:
        net, norm, stats = train_network(sc, a.rules, a.epochs, a.seed, a.label_scenarios,
                                         on_device=not a.labels_on_host)
        _log("network", json.dumps(stats))
        gain = screening_gain(sc, net, norm, a.pool, a.seed + 100)
        _log("screening", json.dumps(gain))
        out.update(network=stats, screening=gain)
        setup_s = sum(stats["seconds"].values())
    if a.generations > 0:
        # warm-up of the device loops (code objects, graph instantiation)
        run_device_gp(sc, None, 2, 64, 64, a.seed + 7, "warm-up")
        if a.online:
            run_device_gp_online(sc, 3, 64, 512, a.seed + 7, "warm-up", train_at=(1, 2))
        else:
            run_device_gp(sc, net, 2, 64, 512, a.seed + 7, "warm-up")
        runs = []
        for i in range(a.gp_runs):
            gseed = a.seed + 7 + i
            plain = run_device_gp(sc, None, a.generations, a.pop, a.pop, gseed, "plain seed %d" % gseed,
                                  launch=launch)
            if a.online:
                surr = run_device_gp_online(sc, a.generations, a.pop, a.breed, gseed, "surrogate seed %d" % gseed,
                                            launch=launch)
            else:
                surr = run_device_gp(sc, net, a.generations, a.pop, a.breed, gseed, "surrogate seed %d" % gseed,
                                     launch=launch, screened_init=a.screened_init)
            target = plain["held_out_best_per_generation"][-1]
            pg, pms = reach(plain["held_out_best_per_generation"], plain["generation_ms"], target)
            sgen, sms = reach(surr["held_out_best_per_generation"], surr["generation_ms"], target)
            total = plain["generation_ms"][-1]
            share = 1000.0 * setup_s / a.gp_runs      # the network's setup, shared by the runs
            runs.append({"seed": gseed, "plain": plain, "surrogate": surr, "target_held_out": target,
                         "plain_reaches_target": {"generation": pg, "device_ms": pms},
                         "surrogate_reaches_target": {"generation": sgen, "device_ms": sms},
                         **ratios_of(sms, pms, total, share)})
            _log("seed %d: target %.2f, plain reaches it at generation %s (%.0f ms), surrogate at %s (%.0f ms); "
                 "plain run %.0f ms, setup share %.0f ms"
                 % (gseed, target, pg, pms, sgen, sms if sms is not None else float("nan"), total, share))
        out["gp"] = {"generations": a.generations, "pop": a.pop, "breed": a.breed, "runs": runs,
                     "screened_init": bool(a.screened_init),
                     "network_setup_s": round(setup_s, 3),
                     "runs_sharing_the_network": 1 if a.online else a.gp_runs}
        for key in ("time_ratio", "time_ratio_with_setup", "time_ratio_to_plain_reach",
                    "time_ratio_to_plain_reach_with_setup"):
            vals = [r[key] if r[key] is not None else float("inf") for r in runs]
            out["gp"]["median_" + key] = float(np.median(vals))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
