"""The C oracle against the reference's golden fixtures (CPU only).

sim.txt pins the simulator (every slot of full61, incl. erf/erfc/pow via
ATRISK and PBUST) and zstats bit for bit; gp.txt pins rmod/pow and 6,144
more rewards."""
import numpy as np
import pytest

from golden_data import bits, gp_cases, sim_cases
from sigp.instance import load_instance
from sigp.rules import compile_tree, pack_trees, parse_prefix, term_mask_from, NS_TERMINALS, EXTRA

CASES = sim_cases()


@pytest.mark.parametrize("case", CASES, ids=[c["inst"] + "-" + c["kind"] for c in CASES])
def test_sim_rewards_and_zstats(oracle, case):
    inst = load_instance(case["inst"])
    zp = oracle.zstats(inst)
    assert np.array_equal(bits(zp), bits(case["zp"])), "zstats differ"
    ops, args = compile_tree(parse_prefix(case["tree"]))
    a = np.array([x for pr in args for x in pr], dtype=np.int32)
    S = case["sampled"]
    got = oracle.sim_batch(inst, ops, a, S, S.shape[0], case["mask"], case["zp"])
    assert np.array_equal(bits(got), bits(case["want"]))


GP = gp_cases()


@pytest.mark.parametrize("ci", range(len(GP)), ids=[c["kind"] for c in GP])
def test_gp_fixture_fitness(oracle, ci):
    """gp.txt fitness values (1 scenario per generation, c101) from the
    fixture's own populations."""
    case = GP[ci]
    inst = load_instance("c101")
    zp = oracle.zstats(inst)
    terms = NS_TERMINALS + (EXTRA if case["kind"] == "ext29rmod" else [])
    mask = term_mask_from(terms)
    for g in case["gens"]:
        pk = pack_trees([parse_prefix(t) for t in g["pop"]])
        got = oracle.evaluate(inst, pk, g["sampled"][None, :], 1, mask, zp, threads=8)
        assert np.array_equal(bits(got), bits(g["fits"]))
