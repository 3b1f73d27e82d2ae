"""A GP loop created for screening (offspring = 4 x population) takes plain
steps while no network is attached: sub-step 0's offspring are the
reference's gp_step on the run's own stream (engine/gp.hpp:170-194), so its
populations and fitness histories equal the plain loop's bit for bit; with
a network attached from a later generation on, the loops part there.  Then
the surrogate arm whose network trains inside the run (sigp.screening.
run_device_gp_online) on a reduced configuration: its record holds every
training point, and its held-out rewards are the exact simulator's."""
import numpy as np
import pytest

from gpu_util import assert_bitwise

pytestmark = pytest.mark.gpu


def _loop(sc, offspring, seed=3, pop=128, gens=4):
    from sigp.gp import EvolutionConfig, TreeConfig
    from sigp.gp_device import DeviceGP
    from sigp.rules import FUNCTIONS
    from sigp.screening import VOCAB
    tc = TreeConfig(list(VOCAB), list(FUNCTIONS), sc.depth, 2)
    ec = EvolutionConfig(population_size=pop, max_depth=sc.depth)
    gp = DeviceGP(sc.ctx, tc, ec, seeds=[seed], T=sc.n_train, term_mask=sc.mask, offspring=offspring)
    gp.init()
    return gp


def test_unscreened_steps_equal_plain_gp():
    from sigp.instance import synthetic_clustered
    from sigp.screening import Scorer
    from sigp.surrogate import Mlp
    sc = Scorer(synthetic_clustered(n=60, m=2, seed=4))
    pop, gens = 128, 4
    rows = np.concatenate([sc.train_scenarios(g) for g in range(gens)])
    plain = _loop(sc, 0)
    screen = _loop(sc, 4 * pop)
    a = plain.run(rows, generations=gens, step_last=True)
    b = screen.run(rows, generations=gens, step_last=True)
    assert_bitwise(b.fitness, a.fitness, "fitness history")
    assert screen.prefixes(0) == plain.prefixes(0)
    # attached from here on, the screen chooses other parents than sub-step 0's
    net = Mlp(seed=5)
    screen.set_surrogate(net, sc.sit)
    rows2 = np.concatenate([sc.train_scenarios(g) for g in range(gens, gens + 2)])
    plain.run(rows2, generations=2, step_last=True)
    screen.run(rows2, generations=2, step_last=True)
    assert screen.prefixes(0) != plain.prefixes(0)
    # detached again, both take plain steps from their own populations
    screen.set_surrogate(None)
    rows3 = sc.train_scenarios(gens + 2)
    c = screen.run(rows3, generations=1)
    assert c.fitness.shape == (1, 1, pop)
    plain.close()
    screen.close()
    net.close()


def test_online_surrogate_run_records_its_training():
    from sigp.instance import synthetic_clustered
    from sigp.screening import Scorer, run_device_gp_online
    sc = Scorer(synthetic_clustered(n=60, m=2, seed=4))
    gens, pop = 8, 128
    res = run_device_gp_online(sc, gens, pop, 4 * pop, 11, "online", train_at=(2, 4, 6))
    assert [t["generation"] for t in res["training"]] == [2, 4, 6]
    assert [t["rows"] for t in res["training"]] == [2 * pop, 4 * pop, 6 * pop]
    ms = res["generation_ms"]
    assert len(ms) == gens and all(b >= a for a, b in zip(ms, ms[1:]))
    assert res["host_ms"] > 0 and ms[-1] > res["host_ms"]
    held = np.asarray(res["held_out_best_per_generation"])
    assert held.shape == (gens,) and np.all(held > 0)


def test_population_phenotype_and_device_label_set():
    """sigp_gp_phenotype of an evolved two-run population equals
    sigp_phenotype of the same trees compiled on the host; the device label
    set's rules are the host's ramped half-and-half of the same seeds and
    its labels their exact means."""
    import random
    from sigp.gp import EvolutionConfig, TreeConfig, ramped_half_and_half
    from sigp.gp_device import DeviceGP
    from sigp.instance import synthetic_clustered
    from sigp.rules import FUNCTIONS, pack_trees
    from sigp.screening import VOCAB, Scorer, device_label_set
    from sigp.surrogate import features, phenotype
    sc = Scorer(synthetic_clustered(n=60, m=2, seed=4))
    tc = TreeConfig(list(VOCAB), list(FUNCTIONS), sc.depth, 2)
    gp = DeviceGP(sc.ctx, tc, EvolutionConfig(population_size=96, max_depth=sc.depth), seeds=[5, 6],
                  T=sc.n_train, term_mask=sc.mask)
    gp.init()
    rows = np.concatenate([np.tile(sc.train_scenarios(g), (2, 1)) for g in range(2)])
    gp.run(rows, generations=2, step_last=True)
    dev = gp.phenotype(sc.sit)
    trees = gp.population(0) + gp.population(1)
    host = phenotype(sc.ctx, pack_trees(trees), sc.mask, sc.sit)
    assert dev.shape == host.shape and np.array_equal(dev, host)
    gp.close()
    scen = np.ascontiguousarray(sc.train_scenarios(0)[:12])
    X, y, _, _ = device_label_set(sc, 300, scen, 7, pop_max=128)   # 3 populations of 100
    assert X.shape == (300, 512) and y.shape == (300,)
    trees = []
    for r in range(3):
        trees += ramped_half_and_half(tc, random.Random(7 * 1_000_003 + r), 100)
    assert_bitwise(y, sc.exact(trees, scen), "labels")
    assert np.array_equal(X, features(phenotype(sc.ctx, pack_trees(trees), sc.mask, sc.sit)))
