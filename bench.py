"""Rule x scenario simulation throughput on one MI355X.

A step evaluates the whole plain-method workload once: 1,024 ramped
half-and-half rules (EXTRA-18 terminals, depth 2-6 levels) on one synthetic
uniform instance (n = 100, m = 2) over 500 common-random-number scenarios
(Gamma(shape 4, mean 10) service times), producing every per-(rule,
scenario) reward and the per-rule means.  Scenarios, rules and the instance
are resident in HBM before timing; each step is one launch of the
simulation kernel (plus the per-rule mean kernel) on torch's current stream,
timed with HIP events on that stream.

Prints one JSON line: value = rule x scenario simulations per second on the
one GPU, with a `roofline` object (the workload's algorithmic FP64
operations per simulation, from the committed work model
tests/golden/work_plain_small.json, times the simulations of one launch,
divided by the simulation kernel's own event-timed duration, against the
MI355X FP64 peak) and a `cpu_baseline` object (the reference's own
simulator, compiled from its sources, timed on the host cores on a bounded
sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "solver-informed-dispatching-rules_amd"))

import numpy as np  # noqa: E402

FP64_PEAK_TFLOPS = 78.6   # MI355X published FP64 (vector and matrix) peak, FMA = 2 flops

# HBM bytes per simulation-kernel launch, keyed by (rules, scenarios): the
# FETCH_SIZE and WRITE_SIZE counters of profiles/sim_kernel_plain_small_hbm_traffic.md,
# each scaled by what it reads for a plain copy of known size (x1.9997, x1.0).
TRAFFIC_BYTES = {(1024, 500): 21430034}
MAPPINGS = {"scenarios": 1, "candidates": 2, "candidates2": 3}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


WORK_MODELS = {"plain": "work_plain_small.json", "large": "work_large.json", "batched": "work_batched.json"}


def work_model(kind, wl):
    """Algorithmic FP64 operations per (rule, scenario) simulation of this
    workload: the committed count of tests/golden/work_<workload>.json (data
    written by tests/golden/make_work_model.py with the counting oracle),
    used only when the workload's arrays match the ones it was counted on."""
    from sigp.workloads import workload_digest
    name = WORK_MODELS.get(kind)
    fn = os.path.join(ROOT, "tests", "golden", name) if name else None
    if fn is None or not os.path.exists(fn):
        return None
    with open(fn) as f:
        doc = json.load(f)
    if doc["digest"] != workload_digest(wl):
        log("work model does not describe this workload: roofline omitted")
        return None
    return doc


def _host_cpu():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"logical_cpus": os.cpu_count(), "model": model}


def cpu_baseline(wl, zp, target_s=15.0, one_core_s=4.0):
    """The reference's simulator (oracle/_ref) on host cores over a bounded
    sample (16 threads, then one thread on a smaller sample); falls back to
    the C restatement when the reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle, Reference
    from sigp.rules import pack_trees
    threads = max(1, min(16, os.cpu_count() or 1))
    if Reference.available():
        impl, kind = Reference(), "reference"
    else:
        impl, kind = Oracle(), "port"

    def run(nr, ns, th):
        step = max(1, wl.packed.n_rules // nr)
        trees = [wl.trees[i] for i in range(0, wl.packed.n_rules, step)][:nr]
        pk = pack_trees(trees)
        S = np.ascontiguousarray(wl.sampled[:ns])
        t0 = time.perf_counter()
        impl.evaluate(wl.inst, pk, S, ns, wl.mask, zp, threads=th)
        return time.perf_counter() - t0, pk.n_rules * ns

    dt, sims = run(threads, 4, threads)
    rate = sims / max(dt, 1e-6)
    want = int(rate * target_s)
    nr = max(threads, min(wl.packed.n_rules, want // 32))
    ns = max(1, min(wl.n_scen, want // nr))
    dt, sims = run(nr, ns, threads)
    # one thread: every rule of the sample above (the rule mix matters more
    # than the scenario count), fewer scenarios
    ns1 = max(1, min(ns, int(sims / dt / threads * one_core_s) // nr))
    dt1, sims1 = run(nr, ns1, 1)
    return {"value": sims / dt, "unit": "sims/s", "cores": threads, "kind": kind,
            "sample": "%d rules (every %d-th) x first %d scenarios of the workload, %.1f s"
                      % (nr, max(1, wl.packed.n_rules // nr), ns, dt),
            "value_1core": sims1 / dt1,
            "sample_1core": "the same rules x first %d scenarios, %.1f s" % (ns1, dt1),
            "host": _host_cpu()}


def describe(kind, wl):
    P, T = wl.packed.n_rules, wl.n_scen
    if kind == "plain":
        return ("plain method, small: 1 uniform instance n=100 m=2, %d rules (ramped half-and-half, "
                "depth 2-6 levels, EXTRA-18 terminals) x %d Gamma(4, mean 10) scenarios" % (P, T))
    if kind == "gp-generation":
        return ("one GP generation: clustered instance n=100 m=3, %d rules (ramped half-and-half, "
                "depth 3-9 levels, NS-11 + EXTRA-18) x %d Gamma(4) scenarios" % (P, T))
    if kind == "large":
        return ("large: clustered + uniform instance n=288 m=4 (windows U[20,200]), %d rules (up to 127 "
                "nodes, EXTRA-18) x %d Gamma(2, mean 10) scenarios" % (P, T))
    if kind == "batched":
        ns = sorted(i.n for i in wl.insts)
        return ("batched uneven: %d instances (n %d-%d, m 1-4, uniform/clustered, windows U[10,200]) "
                "x %d rules (up to 127 nodes, EXTRA-18) x %d Gamma(4) scenarios each, one multi-instance "
                "batch" % (len(wl.insts), ns[0], ns[-1], wl.packed.n_rules, wl.n_scen))
    return ("one launch of the reference GP loop: clustered instance n=100 m=3, %d seeds x %d rules "
            "(NS-11 + EXTRA-18, depth 3-9 levels) x 1 scenario each" % (P // wl.group_size,
                                                                      wl.group_size))


def bench_gp_loop(args):
    """The whole GP loop on the device: clustered instance n=100 m=3, one run
    of 1,024 rules over NS-11 + EXTRA-18 (ramped half-and-half, depth limit
    8, tournament 7, crossover 0.8 / mutation 0.15, 10 elites), 50 training
    scenarios per generation (Gamma 4, seeds run*10000 + generation, as
    engine/gp_engine.cu:69), 50 generations, one generation per HIP-graph
    replay.  A step is one generation: evaluation plus breeding."""
    import torch  # noqa: F401  (one HIP runtime for torch and libsigp)
    from sigp import gp as G
    from sigp.api import FactoryContext, device_info
    from sigp.gp_device import DeviceGP
    from sigp.rules import EXTRA, FUNCTIONS, NS_TERMINALS, pack_trees, term_mask_from
    from sigp.scenarios import gen_durations
    from sigp.workloads import TEST_SEED, clustered_gp_instance

    gens, T, pop, runs = args.generations, args.scenarios or 50, args.rules or 1024, args.runs
    inst = clustered_gp_instance()
    vocab = list(NS_TERMINALS) + list(EXTRA)
    mask = term_mask_from(vocab)
    tc = G.TreeConfig(vocab, list(FUNCTIONS))
    ec = G.EvolutionConfig(population_size=pop)
    t0 = time.perf_counter()
    rows = np.concatenate([gen_durations(inst.durations, T, r * 10000 + g, mode="gamma", shape=4.0)
                           for g in range(gens) for r in range(runs)])
    log("scenarios for %d generations built in %.1fs" % (gens, time.perf_counter() - t0))
    name, n_cu, _ = device_info()
    log("device", name, n_cu, "CUs")
    ctx = FactoryContext(inst)
    seeds = list(range(runs))
    launch = {}
    if args.waves:
        launch = dict(waves_per_item=args.waves, mapping=MAPPINGS[args.mapping])
    # warm-up: code objects, graph instantiation
    gp = DeviceGP(ctx, tc, ec, seeds=seeds, T=T, term_mask=mask, **launch)
    gp.init()
    gp.run(rows[:2 * runs * T], generations=2, graph=True)
    gp.close()
    gp = DeviceGP(ctx, tc, ec, seeds=seeds, T=T, term_mask=mask, **launch)
    gp.init()
    wall0 = time.perf_counter()
    res = gp.run(rows, generations=gens, graph=not args.no_graph)
    hist, dev_ms = res.fitness, res.ms
    wall = time.perf_counter() - wall0
    log("device loop: %d generations in %.1f ms device time, %.3f s wall" % (gens, dev_ms, wall))
    # the reference's read-out: best of the final population on one more
    # training draw, then its mean over 500 held-out scenarios
    final = gp.population()
    best_rules = []
    for r in range(runs):
        fs = gen_durations(inst.durations, T, r * 10000 + gens, mode="gamma", shape=4.0)
        _, fm = ctx.evaluate_means(pack_trees(final[r]), fs, T, mask)
        best_rules.append(final[r][int(np.argmax(fm))])
    test = gen_durations(inst.durations, 500, TEST_SEED, mode="gamma", shape=4.0)
    _, held = ctx.evaluate_means(pack_trees(best_rules), test, 500, mask)
    log("held-out 500-scenario mean reward of each run's best rule:", np.round(held, 3))
    sims = gens * runs * pop * T
    ms_gen = dev_ms / gens
    out = {
        "metric": "rule x scenario sims/sec (whole GP loop: evaluation + breeding on the device)",
        "value": round(sims / (dev_ms / 1000.0), 1), "unit": "sims/s", "n_gpus": 1, "steps": gens,
        "warmup": 2, "ms_per_step": round(ms_gen, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "GP loop on the device: clustered instance n=100 m=3 (windows U[30,120]), "
                               "%d run(s) x %d rules (NS-11 + EXTRA-18, depth <= 8), %d Gamma(4) training "
                               "scenarios per generation, %d generations, %s; value counts every tree of "
                               "each generation, of which each distinct tree is simulated once (repeats "
                               "copy its rewards, bit for bit)"
                               % (runs, pop, T, gens, "one HIP graph per generation" if not args.no_graph
                                  else "stream launches"),
                   "rules": pop * runs, "scenarios": T, "generations": gens, "n": inst.n, "m": inst.m,
                   "parallelism": "1 GPU; evaluation lanes: %s" % (
                       "automatic" if not args.waves else "%s, %d wavefronts" % (args.mapping, args.waves))},
        "wall_s": round(wall, 3),
        "held_out_reward": [round(float(v), 3) for v in held],
        "best_train_fitness_per_generation": [round(float(v), 3) for v in hist[:, 0].max(axis=1)],
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


# The reference's published training times (results/timing/timing_bench.log:1-6,
# its own GPU, unnamed): `timing <instance> [--terms ...] --seeds 30`.
TIMING_PUBLISHED = {("c101", ""): 21.0, ("c101", "NSTW,MAXN,NFEAS,REACH5"): 36.0,
                    ("r201", ""): 69.0, ("r201", "NSTW,MAXN,NFEAS,REACH5"): 134.0,
                    ("pr15", ""): 188.0, ("pr15", "NSTW,MAXN,NFEAS,REACH5"): 1092.0}


def bench_timing(args):
    """s3_generation/timing.cu on the device: 30 seeds of GP (population
    1,024, 51 generations, one training scenario per generation, NS-11 plus
    --terms) trained side by side with the loop on the device, then each
    seed's final rule (first maximum on one more training draw) scored on
    500 held-out scenarios.  The timed training covers what the reference's
    does: every generation's scenario rows (drawn on the host, as there),
    evaluation and GP step, and the final draw."""
    import torch
    from sigp import gp as G
    from sigp.api import FactoryContext, device_info
    from sigp.gp_device import DeviceGP, tokens_to_tree
    from sigp.instance import load_instance
    from sigp.rules import FUNCTIONS, NS_TERMINALS, pack_trees, term_mask_from
    from sigp.scenarios import gen_durations

    name, terms = args.instance, args.terms
    extra = [t for t in terms.split(",") if t]
    inst = load_instance(name)
    ctx = FactoryContext(inst)
    vocab = list(NS_TERMINALS) + extra
    mask = term_mask_from(vocab)
    tc, ec = G.TreeConfig(vocab, list(FUNCTIONS)), G.EvolutionConfig()
    seeds, T, gens = list(range(30)), ec.train_samples, ec.generations
    log("device", device_info()[0])

    def train():
        gp = DeviceGP(ctx, tc, ec, seeds=seeds, T=T, term_mask=mask)
        gp.init()
        rows = np.concatenate([gen_durations(inst.durations, T, s * 10000 + g) for g in range(gens) for s in seeds])
        gp.run(rows, generations=gens)
        fit = gp.evaluate(np.concatenate([gen_durations(inst.durations, T, s * 10000 + gens) for s in seeds]))
        tok, ln = gp.tokens()
        rules = [tokens_to_tree(tok[s, int(np.argmax(fit[s])), :ln[s, int(np.argmax(fit[s]))]], tc.terminals,
                                tc.functions) for s in seeds]
        gp.close()
        return rules

    tw = time.perf_counter()
    train()                                   # warm-up: code objects, graph instantiation
    torch.cuda.synchronize()
    log("warm-up training done in %.1f s" % (time.perf_counter() - tw))
    t0 = time.perf_counter()
    rules = train()
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    test = gen_durations(inst.durations, 500, 1000000)
    t1 = time.perf_counter()
    _, held = ctx.evaluate_means(pack_trees(rules), test, 500, mask)
    held_s = time.perf_counter() - t1
    log("TIMING %s [%s] vocab=%d seeds=30 | train %.2fs | heldout-eval %.3fs"
        % (name, terms or "NSGP-baseline-11", len(vocab), train_s, held_s))
    ref = TIMING_PUBLISHED.get((name, terms))
    out = {
        "metric": "GP training time, 30 seeds x 51 generations x 1,024 rules (timing.cu)",
        "value": round(train_s, 3), "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(train_s * 1000, 1), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None if ref is None else round(ref / train_s, 2), "dtype": "f64",
        "data": "reference instance %s (the reference's own file), service times from the reference's "
                "seeded streams" % name,
        "config": {"workload": "timing %s%s --seeds 30: the reference's training on the device (same "
                               "populations as its host loop)" % (name, (" --terms " + terms) if terms else ""),
                   "instance": name, "vocab": len(vocab), "parallelism": "1 GPU; 30 seeds side by side"},
        "reference_published_s": ref, "heldout_eval_s": round(held_s, 4),
        "heldout_mean_reward": round(float(np.mean(held)), 3),
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


# The GA's 18-terminal universe, in the order of the reference's runs
# (results/ga40/run_ga40.sh); bit b of a genome selects UNIVERSE[b].
UNIVERSE = ["REGRET", "NSROB", "MAXN", "REACH2", "NSTW", "FRAGCNT", "FRAGSCORE", "PBUST", "REACH1", "REACH3",
            "REACH5", "SAT", "VFRAC", "NFEAS", "HARVEST", "BOLSAO", "ATRISK", "REACH2ROB"]
# Logged times of the reference's tools (its GPU, unnamed): the GA's first
# batch (results/ga40/logs/c101_search.log:2) and gen_trace with all 18
# extra terminals (results/ga40/logs_full18/full18_queue.log).
GA_BATCH_LOGGED = {"c101": 563.9}
TRACE_FULL18_LOGGED = {"pr01": 84.0, "c101": 154.0}


def _timed_line(metric, value, ref, workload, parallelism, instance, extra):
    out = {"metric": metric, "value": round(value, 3), "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 1,
           "ms_per_step": round(value * 1000, 1), "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None if ref is None else round(ref / value, 2), "dtype": "f64",
           "data": "reference instance %s (the reference's own file), service times from the "
                   "reference's seeded streams" % instance,
           "config": {"workload": workload, "parallelism": parallelism},
           "reference_logged_s": ref, "roofline": None, "cpu_baseline": None}
    out.update(extra)
    print(json.dumps(out), flush=True)


def bench_ga_batch(args):
    """The GA's first batch on the device: 13 genomes (the NS-11 base and the
    first 12 single-terminal subsets of the 18-terminal universe) x 16 seeds,
    each the proxy fitness of s4_selection/ga.cu:109-198 (51 generations of
    GP from Mt19937(seed), then the final rule on the 64-scenario train
    block), with every seed's GP loop on the device."""
    import torch
    from sigp.api import FactoryContext, device_info
    from sigp.ga_device import proxy_scores
    from sigp.instance import load_instance
    ctx = FactoryContext(load_instance(args.instance))
    log("device", device_info()[0])
    proxy_scores(ctx, UNIVERSE, 0, proxy=2, pgens=2)            # warm-up
    torch.cuda.synchronize()
    genomes = [0] + [1 << b for b in range(12)]
    t0 = time.perf_counter()
    for g in genomes:
        proxy_scores(ctx, UNIVERSE, g)
        log("genome %#x done at %.1f s" % (g, time.perf_counter() - t0))
    dt = time.perf_counter() - t0
    _timed_line("GA proxy evaluation time, 13 genomes x 16 seeds x 51 generations (ga.cu batch_score)", dt,
                GA_BATCH_LOGGED.get(args.instance),
                "ga %s first batch on the device: 13 genomes x 16 seeds, 18-terminal universe" % args.instance,
                "1 GPU; genomes one after another, each genome's 16 seeds' GP loops side by side (the "
                "reference ran the 13 x 16 runs as one lockstep batch)", args.instance,
                {"instance": args.instance})


# The reference's logged whole GA search on c204 (results/ga40/logs/
# c204_search.log): 85 evaluations (80 at 16 seeds, 5 confirmations at 30)
# in 30,236 s of wall time on its unnamed GPU, resumed from 26 cached ones.
GA_SEARCH_LOGGED = {"c204": (30236.0, 85)}


def bench_ga_search(args):
    """The GA terminal-subset search (s4_selection/ga.cu:267-520, restated in
    sigp/ga_search.py) with the proxy fitness of every genome trained by the
    device GP loop: neutral seeds, 13-genome tree, generations until
    stagnation (10), top-5 finalists confirmed at 30 seeds.  Starts from an
    empty cache, so every genome the search visits is trained here."""
    import torch
    from sigp.api import FactoryContext, device_info
    from sigp.ga_device import proxy_batch
    from sigp.ga_search import ga_search
    from sigp.instance import load_instance
    ctx = FactoryContext(load_instance(args.instance))
    log("device", device_info()[0])
    from sigp.ga_device import proxy_scores
    proxy_scores(ctx, UNIVERSE, 0, proxy=2, pgens=2)            # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = ga_search(UNIVERSE, proxy_batch(ctx, UNIVERSE), proxy_batch(ctx, UNIVERSE, proxy=30),
                    seed_mode="neutral", nofull=True, pm=0.5, gens=50, stag_k=10, top_n=5,
                    log=lambda s: log("%7.1f s  %s" % (time.perf_counter() - t0, s)))
    dt = time.perf_counter() - t0
    ref = GA_SEARCH_LOGGED.get(args.instance)
    per_eval_ref = ref[0] / ref[1] if ref else None
    _timed_line("GA terminal-subset search wall time (ga.cu, all proxy evaluations on the device)", dt,
                None, "ga %s --seeds neutral --nofull --pm 0.5 --gens 50 --stag 10 --top 5 on the device, "
                "empty cache" % args.instance,
                "1 GPU; genomes one after another, each genome's 16 (confirmation: 30) seeds' GP loops side by "
                "side", args.instance,
                {"instance": args.instance, "evals": res.evals, "s_per_eval": round(dt / max(1, res.evals), 3),
                 "reference_logged_s_per_eval": None if per_eval_ref is None else round(per_eval_ref, 1),
                 "speedup_per_eval": None if per_eval_ref is None else round(per_eval_ref / (dt / res.evals), 1),
                 "champion": res.lines[-1]})


def bench_trace(args):
    """gen_trace <instance> --terms <all 18> --seeds 30 on the device
    (sigp/trace_device.py): 30 seeds x 51 generations of GP side by side,
    per-generation records, finals on 500 held-out scenarios."""
    import csv
    import torch
    from sigp.api import FactoryContext, device_info
    from sigp.instance import load_instance
    from sigp.trace_device import trace
    with open(os.path.join(ROOT, "tests", "golden", "cap2-results.csv")) as f:
        km = {r["instance"]: float(r["baseline_nsgp_mean30"]) for r in csv.DictReader(f)}[args.instance]
    terms = [t for t in args.terms.split(",") if t] or UNIVERSE
    ctx = FactoryContext(load_instance(args.instance))
    log("device", device_info()[0])
    trace(ctx, args.instance, terms, km, n_seeds=1)             # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = trace(ctx, args.instance, terms, km)
    dt = time.perf_counter() - t0
    log("trace %s (%d extra terminals): %.2f s, mean gain %.2f" % (args.instance, len(terms), dt, res["mean_gain"]))
    ref = TRACE_FULL18_LOGGED.get(args.instance) if terms == UNIVERSE else None
    _timed_line("GP trace time, 30 seeds x 51 generations (gen_trace.cu)", dt, ref,
                "gen_trace %s --terms %s --seeds 30 on the device" % (args.instance, ",".join(terms)),
                "1 GPU; the 30 seeds' GP loops side by side", args.instance,
                {"instance": args.instance, "mean_gain": res["mean_gain"]})


BF16_MFMA_PEAK_TFLOPS = 2500.0   # MI355X dense BF16 matrix peak
MLP_FWD_FLOP = 2 * (512 * 256 + 256 * 256)          # the two matrix layers, per sample
MLP_TRAIN_FLOP = MLP_FWD_FLOP + 2 * (256 * 256 + 256 * 256 + 512 * 256)   # + delta1, dW2, dW1


def bench_mlp(args):
    """The surrogate network on the matrix cores: training on 100,000
    phenotype-shaped rows (512 ranks / 7 on an 8-level grid) for 20 epochs at
    batch 256 (the screen's own training set-up), then scoring 8,192 rows
    (one generation's screen).  FLOPs count the matrix products only, per
    sample: 393,216 forward, 917,504 for a training step."""
    import torch
    from sigp.api import device_info
    from sigp.surrogate import Mlp
    rows, epochs = args.rules or 100_000, 20
    rng = np.random.default_rng(1)
    x = (rng.integers(0, 8, (rows, 512)) / 7.0).astype(np.float32)
    w = rng.normal(size=512) / np.sqrt(512)
    t = np.tanh(x @ w).astype(np.float32)
    log("device", device_info()[0])
    net = Mlp(seed=1)
    net.train(x[:4096], t[:4096], epochs=1, batch=256)      # warm-up (code objects, buffers)
    net.train(x, t, epochs=1, batch=256)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = net.train(x, t, epochs=epochs, batch=256, lr=1e-3, seed=3)
    train_s = time.perf_counter() - t0
    xs = x[:8192]
    net.predict(xs)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        net.predict(xs)
    pred_s = (time.perf_counter() - t0) / args.steps
    samples = rows * epochs
    achieved = MLP_TRAIN_FLOP * samples / train_s / 1e12
    log("training %.3f s (%.1f us per step), losses %.4f -> %.4f; scoring 8,192 rows %.2f ms"
        % (train_s, train_s / (samples / 256) * 1e6, losses[0], losses[-1], pred_s * 1e3))
    out = {"metric": "surrogate training samples/sec", "value": round(samples / train_s, 1), "unit": "samples/s",
           "n_gpus": 1, "steps": epochs, "warmup": 1, "ms_per_step": round(train_s / epochs * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
           "data": "synthetic phenotype-shaped rows",
           "config": {"workload": "MLP 512-256-256-1, %d rows x %d epochs, batch 256, Adam lr 1e-3; then "
                                  "scoring 8,192 rows" % (rows, epochs),
                      "parallelism": "1 GPU; one wavefront per 16 x 16 output tile"},
           "roofline": {"bound": "mfma", "achieved": round(achieved, 4), "peak": BF16_MFMA_PEAK_TFLOPS,
                        "unit": "TFLOP/s", "frac": round(achieved / BF16_MFMA_PEAK_TFLOPS, 6), "traffic": None,
                        "flop_per_sample": MLP_TRAIN_FLOP, "note": "whole training wall time, host shuffle, "
                        "upload and per-step launches included"},
           "cpu_baseline": None, "train_s": round(train_s, 3), "score_8192_ms": round(pred_s * 1e3, 3),
           "score_tflops": round(MLP_FWD_FLOP * 8192 / pred_s / 1e12, 4)}
    print(json.dumps(out), flush=True)


def bench_sampler(args):
    """Scenario rows on the device: the large configuration's 2,000 Gamma(2)
    rows of n = 288 (clustered + uniform, m = 4), each row its own
    Mt19937(1,000,000 + r) stream.  A step is one full set of rows; the
    reference samples them on the host (engine/sim_host.hpp:50-58), timed
    beside it on the same rows with the host restatement."""
    import torch
    from sigp.api import FactoryContext, device_info, sampler
    from sigp._lib import check, lib
    from sigp.instance import synthetic_mixed
    from sigp.scenarios import gen_durations
    import ctypes as C
    rows = args.scenarios or 2000
    inst = synthetic_mixed(n=288, m=4, seed=1)
    ctx = FactoryContext(inst)
    name, n_cu, _ = device_info()
    log("device", name, n_cu, "CUs")
    smp = sampler("gamma", 0.2, 2.0)
    out = torch.zeros(rows * inst.n, dtype=torch.float64, device="cuda:0")
    stream = torch.cuda.current_stream()

    def step():
        check(lib().sigp_sample_durations_async(ctx._h, C.byref(smp), 1_000_000, rows,
                                                C.c_void_p(out.data_ptr()), C.c_void_p(stream.cuda_stream)))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    for a, b in evs:
        a.record(stream)
        step()
        b.record(stream)
    torch.cuda.synchronize()
    ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))
    sample = min(rows, 100)
    t0 = time.perf_counter()
    gen_durations(inst.durations, sample, 1_000_000, mode="gamma", shape=2.0)
    host_s = time.perf_counter() - t0
    out_d = {"metric": "scenario rows/sec (device sampler)", "value": round(rows / (ms / 1000.0), 1),
             "unit": "rows/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
             "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
             "dtype": "f64", "data": "synthetic",
             "config": {"workload": "device sampler: %d Gamma(2, mean 10) rows of n=%d (clustered + uniform "
                                    "instance, m=4), one Mersenne Twister stream per row" % (rows, inst.n),
                        "rows": rows, "n": inst.n, "parallelism": "1 GPU; one wavefront per row"},
             "roofline": None,
             "cpu_baseline": {"value": round(sample / host_s, 1), "unit": "rows/s", "cores": 1, "kind": "port",
                              "sample": "the first %d rows with the host sampler (Python, libm)" % sample}}
    print(json.dumps(out_d), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--rules", type=int, default=None, help="rules (plain: 1,024; large: 4,096)")
    ap.add_argument("--scenarios", type=int, default=None, help="scenarios (plain: 500; large: 2,000)")
    ap.add_argument("--workload", choices=["plain", "gp-generation", "gp-reference", "batched", "large",
                                           "gp-loop", "sampler", "timing", "ga-batch", "ga-search", "trace",
                                           "mlp"],
                    default="plain",
                    help="plain: the headline workload; gp-generation: one generation of the "
                         "on-device GP loop (1,024 rules x 50 scenarios, clustered n=100 m=3); "
                         "gp-reference: one launch of the reference's GP loop (30 seeds x 1,024 "
                         "rules x 1 scenario); batched: 40 uneven instances (n 48-288) x 256 "
                         "rules x 500 scenarios in one multi-instance batch; large: n=288 m=4, "
                         "4,096 rules x 2,000 scenarios; gp-loop: the whole GP loop on the "
                         "device (1,024 rules x 50 scenarios x 50 generations, breeding included); "
                         "sampler: 2,000 Gamma(2) scenario rows of n=288 sampled on the device; "
                         "timing: the reference's timing tool (30 seeds of GP) on the device, "
                         "against its published training times; ga-batch: the GA's first batch "
                         "of 13 genomes x 16 seeds (--instance c101 has a logged time); ga-search: the "
                         "whole GA terminal-subset search (--instance c204 has a logged time); trace: "
                         "gen_trace with all 18 extra terminals (--instance pr01 or c101 logged); mlp: "
                         "the surrogate network's training (100,000 rows x 20 epochs) and scoring on "
                         "the matrix cores")
    ap.add_argument("--generations", type=int, default=50, help="gp-loop: generations")
    ap.add_argument("--instance", default="c101", help="timing: instance (c101, r201, pr15 are published)")
    ap.add_argument("--terms", default="",
                    help="timing: extra terminals beside NS-11, e.g. NSTW,MAXN,NFEAS,REACH5 (empty: NS-11 "
                         "alone); trace: the extra terminals of the trace (empty: all 18 of the GA's "
                         "universe)")
    ap.add_argument("--runs", type=int, default=1, help="gp-loop: independent runs side by side")
    ap.add_argument("--no-graph", action="store_true", help="gp-loop: stream launches, no HIP graph")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--waves", type=int, default=0,
                    help="wavefronts per workgroup: 1, 2 or 4; 0 = autotune")
    ap.add_argument("--mapping", choices=sorted(MAPPINGS), default="scenarios",
                    help="lane mapping used with --waves (candidates2: 2 simulations per wavefront, "
                         "--waves 1)")
    ap.add_argument("--protocol", choices=["vehicles", "rr"], default="vehicles",
                    help="plain / large / batched: vehicles = sim_one (the reference's V0, the "
                         "headline); rr = sim_one_rr, the round-robin ablation arm (no roofline: the "
                         "committed work models count V0)")
    ap.add_argument("--one-launch", action="store_true",
                    help="batched: every size class in one launch (sigp_instance_one_launch) instead of one "
                         "launch per class on concurrent streams")
    ap.add_argument("--no-scenario-order", action="store_true",
                    help="scenario lanes in index order (sigp_batch_order_scenarios off)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after timing, run one device copy of this many MiB (PMC calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(each GPU would evaluate its own share of rules); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "gp-loop":
        return bench_gp_loop(args)
    if args.workload == "sampler":
        return bench_sampler(args)
    if args.workload == "timing":
        return bench_timing(args)
    if args.workload == "ga-batch":
        return bench_ga_batch(args)
    if args.workload == "ga-search":
        return bench_ga_search(args)
    if args.workload == "trace":
        return bench_trace(args)
    if args.workload == "mlp":
        return bench_mlp(args)
    import torch
    from sigp.api import FactoryContext, batch_multi, device_info
    from sigp.workloads import (batched_uneven, gp_generation, gp_reference_generation, large,
                                plain_small)

    t0 = time.perf_counter()
    if args.workload == "plain":
        wl = plain_small(n_rules=args.rules or 1024, n_scen=args.scenarios or 500)
    elif args.workload == "gp-generation":
        wl = gp_generation()
    elif args.workload == "gp-reference":
        wl = gp_reference_generation()
    elif args.workload == "large":
        wl = large(n_rules=args.rules or 4096, n_scen=args.scenarios or 2000)
    else:
        wl = batched_uneven()
    log("workload built in %.1fs: %s" % (time.perf_counter() - t0, wl.name))
    dev = torch.device("cuda:0")
    name, n_cu, clk = device_info()
    log("device", name, n_cu, "CUs")
    T = wl.n_scen
    rr = args.protocol == "rr"
    if args.workload == "batched":
        ctxs = [FactoryContext(i) for i in wl.insts]
        for c in ctxs:
            c.protocol(rr)
            c.one_launch(args.one_launch)
        ctx = ctxs[0]
        batch = batch_multi(ctxs, wl.packed, T, wl.mask)
        n_rows = len(wl.insts) * T
        host_rows = np.concatenate([b.ravel() for b in wl.blocks])
        P = len(wl.insts) * wl.packed.n_rules     # (instance, rule) pairs
    else:
        ctx = FactoryContext(wl.inst)
        ctx.protocol(rr)
        batch = ctx.batch(wl.packed, T, wl.mask, group_size=wl.group_size or None)
        n_rows = wl.sampled.shape[0]
        host_rows = wl.sampled
        P = wl.packed.n_rules
    ordered = batch.order_scenarios(not args.no_scenario_order)
    log("scenario order for scenario lanes:", "on" if ordered else "off")
    sampled = torch.from_numpy(host_rows).to(dev)
    pairs = torch.zeros(P * T, dtype=torch.float64, device=dev)
    means = torch.zeros(P, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)

    def step():
        batch.run(sampled.data_ptr(), n_rows, pairs.data_ptr(), means.data_ptr(), stream.cuda_stream)

    if args.waves:
        mapping = MAPPINGS[args.mapping]
        batch.configure(args.waves, mapping)
        waves, how = args.waves, "fixed"
    else:
        log("autotuning the launch mapping ...")
        (waves, mapping), how = batch.autotune(sampled.data_ptr(), n_rows, stream.cuda_stream), "autotuned"
    lanes = {1: "64-scenario block", 2: "(rule, scenario) simulation", 3: "wavefront of 2 simulations"}[mapping]
    log("%s mapping: %d wavefronts per %s" % (how, waves, lanes))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    for a, b in evs:
        a.record(stream)
        step()
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms = [a.elapsed_time(b) for a, b in evs]
    ms_step = float(np.mean(ms))
    log("per-step ms:", ["%.2f" % v for v in ms], "wall %.3fs" % wall)
    sims = P * T
    value = sims / (ms_step / 1000.0)
    # determinism / sanity of the timed results
    pr = pairs.cpu().numpy()
    assert np.all(np.isfinite(pr)) and np.all(pr == np.round(pr))

    roof = None
    # the work model and the reference's CPU timing exist for the headline workload
    model = None if (args.no_roofline or rr) else work_model(args.workload, wl)
    if model is not None:
        # The simulation kernel(s) alone (no per-rule mean kernel), same
        # stream; a multi-instance batch forks its size classes onto side
        # streams and joins them back, so the events bracket all of them.
        kev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(args.steps)]
        for a, b in kev:
            a.record(stream)
            batch.run(sampled.data_ptr(), n_rows, pairs.data_ptr(), None, stream.cuda_stream)
            b.record(stream)
        torch.cuda.synchronize()
        k_ms = float(np.mean([a.elapsed_time(b) for a, b in kev]))
        log("simulation kernel ms:", "%.3f" % k_ms)
        ops_per_sim = model["ops_per_sim"]
        achieved = ops_per_sim * sims / (k_ms / 1000.0) / 1e12
        roof = {"bound": "valu-fp64", "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 5),
                "traffic": TRAFFIC_BYTES.get((P, T)),
                "kernel_ms": round(k_ms, 3), "ops_per_sim": round(ops_per_sim, 1),
                "issue_frac": round(achieved / (FP64_PEAK_TFLOPS / 2), 5)}
    if args.copy_probe_mib > 0:
        # One plain device-to-device copy of known size, so that a PMC run of
        # this command can scale FETCH_SIZE / WRITE_SIZE to bytes.
        nb = args.copy_probe_mib << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy probe: %d bytes read, %d bytes written" % (nb, nb))
        del src, dst
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(wl, ctx.zp) if (args.workload == "plain" and not rr) else None
        log("cpu baseline", cpu)
    out = {
        "metric": "rule x scenario sims/sec", "value": round(value, 1), "unit": "sims/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": describe(args.workload, wl) + ("; round-robin protocol (sim_one_rr)" if rr else "")
                   + ("; every size class in one launch" if args.one_launch and args.workload == "batched" else ""),
                   "rules": P, "scenarios": T,
                   "n": wl.inst.n if hasattr(wl, "inst") else None,
                   "m": wl.inst.m if hasattr(wl, "inst") else None,
                   "parallelism": "1 GPU; workgroup = %s, %d wavefronts (%s)"
                                  % ({1: "rule x 64 scenarios", 2: "one (rule, scenario) simulation",
                                      3: "2 (rule, scenario) simulations"}[mapping], waves, how)},
        "roofline": roof, "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
