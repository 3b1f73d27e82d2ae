"""The GA subset search's proxy fitness with its inner GP loop on the device
(sigp/ga_device.py) equals the reference's stored evaluation cache
(results/ga40/logs/c204.evcache, tests/golden/gacache): the 16 per-seed
train-block means of each genome, as hexfloats, bit for bit.  Genomes of the
GA's initial pool: the NS-11 base and single-terminal subsets, except
{ATRISK}, whose stored values come from the reference's CUDA erf."""
import gzip
import os
import time

import pytest

from sigp.instance import load_instance

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
UNIVERSE = ["REGRET", "NSROB", "MAXN", "REACH2", "NSTW", "FRAGCNT", "FRAGSCORE", "PBUST", "REACH1", "REACH3",
            "REACH5", "SAT", "VFRAC", "NFEAS", "HARVEST", "BOLSAO", "ATRISK", "REACH2ROB"]


def _cache():
    P = {}
    for ln in gzip.open(os.path.join(GOLDEN, "gacache", "c204.evcache.gz"), "rt"):
        p = ln.split()
        if p and p[0] == "P":
            P[int(p[1], 16)] = [float.fromhex(x) for x in p[3:]]
    return P


@pytest.mark.parametrize("genome", [0, 1 << 0, 1 << 4, 1 << 9, 1 << 13, 1 << 17],
                         ids=["base", "REGRET", "NSTW", "REACH3", "NFEAS", "REACH2ROB"])
def test_device_proxy_equals_the_stored_cache(genome):
    from sigp.api import FactoryContext
    from sigp.ga_device import proxy_scores
    want = _cache()[genome]
    ctx = FactoryContext(load_instance("c204"))
    t0 = time.perf_counter()
    got = proxy_scores(ctx, UNIVERSE, genome)
    print("genome %#x: 16 seeds x 51 generations in %.2f s" % (genome, time.perf_counter() - t0), flush=True)
    assert len(want) == 16
    assert [v.hex() for v in got] == [v.hex() for v in want]
