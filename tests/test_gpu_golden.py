"""GPU path (C ABI) against the reference's golden fixtures, bit for bit."""
import numpy as np
import pytest

from golden_data import bits, gp_cases, sim_cases
from gpu_util import assert_bitwise
from sigp.instance import load_instance
from sigp.rules import EXTRA, NS_TERMINALS, pack_trees, parse_prefix, term_mask_from

pytestmark = pytest.mark.gpu
CASES = sim_cases()


@pytest.mark.parametrize("case", CASES, ids=[c["inst"] + "-" + c["kind"] for c in CASES])
def test_sim_fixture_on_gpu(case):
    from sigp.api import FactoryContext, zp_of
    inst = load_instance(case["inst"])
    print("case", case["inst"], case["kind"], flush=True)
    assert_bitwise(zp_of(inst), case["zp"], "zstats")
    ctx = FactoryContext(inst, case["zp"])
    pk = pack_trees([parse_prefix(case["tree"])])
    S = case["sampled"]
    got = ctx.evaluate(pk, S, S.shape[0], case["mask"])
    assert_bitwise(got, case["want"], "rewards")


GP = gp_cases()


@pytest.mark.parametrize("ci", range(len(GP)), ids=[c["kind"] for c in GP])
def test_gp_fixture_on_gpu(ci):
    """6 x 1,024 fitness values of gp.txt (ext29rmod exercises rmod/pow)."""
    from sigp.api import FactoryContext, zp_of
    case = GP[ci]
    inst = load_instance("c101")
    ctx = FactoryContext(inst)
    terms = NS_TERMINALS + (EXTRA if case["kind"] == "ext29rmod" else [])
    mask = term_mask_from(terms)
    for gi, g in enumerate(case["gens"]):
        print("gp", case["kind"], "gen", gi, flush=True)
        pk = pack_trees([parse_prefix(t) for t in g["pop"]])
        got = ctx.evaluate(pk, g["sampled"][None, :], 1, mask)
        assert_bitwise(got, g["fits"], "gp fitness gen %d" % gi)
