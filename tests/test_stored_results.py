"""The oracle against the reference's stored per-rule results
(tests/golden/stored_results.json.gz, written by make_stored_results.py from
results/baseline/nsgp/<inst>/seed<k>.json): each NS-GP baseline seed's final
rule, re-scored over the 500 held-out scenarios seeded 1,000,000 + s
(s4_selection/redeploy.cu:14-15, engine/sim_host.hpp:50-58) with the NS-11
terminal mask, gives the stored `test_mean` to the digits it was stored
with.  One instance per family here (the GPU library is checked on all 40
instances x 30 seeds in test_gpu_stored_results.py)."""
import gzip
import json
import os

import numpy as np
import pytest

from sigp.instance import load_instance
from sigp.rules import NS_TERMINALS, pack_trees, parse_prefix, term_mask_from
from sigp.scenarios import gen_durations

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TEST_SEED, N_TEST = 1000000, 500


def stored():
    return json.load(gzip.open(os.path.join(GOLDEN, "stored_results.json.gz"), "rt"))


def within_stored_digits(got, text):
    """|got - stored| within half a unit of the stored text's last digit."""
    dec = len(text.split(".")[1]) if "." in text else 0
    return abs(got - float(text)) <= 0.5 * 10.0 ** -dec + 1e-9


def nsgp_case(name):
    seeds = stored()["nsgp"][name]
    inst = load_instance(name)
    trees = [parse_prefix(t) for t, _ in seeds]
    return inst, pack_trees(trees), [w for _, w in seeds]


@pytest.mark.parametrize("name", ["c101", "c201", "r102", "r205", "rc103", "rc202", "pr02"])
def test_oracle_reproduces_stored_nsgp_test_means(oracle, name):
    inst, pk, want = nsgp_case(name)
    test = gen_durations(inst.durations, N_TEST, TEST_SEED)
    zp = oracle.zstats(inst)
    got = oracle.evaluate(inst, pk, test, N_TEST, term_mask_from(NS_TERMINALS), zp,
                          threads=min(8, os.cpu_count() or 1)).reshape(pk.n_rules, N_TEST).mean(axis=1)
    bad = [(k, float(g), w) for k, (g, w) in enumerate(zip(got, want)) if not within_stored_digits(g, w)]
    assert not bad, bad[:5]


def test_fixture_covers_the_stored_runs():
    d = stored()
    assert len(d["nsgp"]) == 40 and all(len(v) == 30 for v in d["nsgp"].values())
    assert len(d["traces"]) == 51
    for t in d["traces"].values():
        assert len(t["finals"]) == 30
        assert np.isfinite(t["mean_gain"])
