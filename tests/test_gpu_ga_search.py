"""The GA terminal-subset search (sigp/ga_search.py) with its proxy fitness
computed by the device GP loop (sigp/ga_device.py), against the reference's
stored c204 run (tests/golden/gasearch): started from the 26 cache entries
the stored run had loaded, three generations evaluate 15 new genomes (none
holding ATRISK, whose stored values come from the reference's CUDA build);
their per-seed proxy values must equal the stored cache's hexfloats bit for
bit, and the generation lines the stored log's."""
import gzip
import os

import pytest

from test_ga_search import KEEP, _options, _stored

pytestmark = pytest.mark.gpu


def test_device_search_reproduces_first_generations():
    from sigp.api import FactoryContext
    from sigp.ga_device import proxy_batch
    from sigp.ga_search import ga_search
    from sigp.instance import load_instance
    log, header, entries = _stored("c204")
    opts, loaded = _options(log, header)
    universe = opts.pop("universe")
    opts.update(gens=3, top_n=1)
    P = {g: v for k, g, v in entries if k == "P"}
    ctx = FactoryContext(load_instance("c204"))
    dev = proxy_batch(ctx, universe)
    done = []

    def proxy(gs):
        print("device proxy for %d genomes" % len(gs), flush=True)
        out = dev(gs)
        done.extend(zip(gs, out))
        return out

    conf = proxy_batch(ctx, universe, proxy=30, pgens=51)
    res = ga_search(universe, proxy, conf, cache={g: v for k, g, v in entries[:loaded] if k == "P"},
                    log=lambda s: print(s, flush=True), **opts)
    assert len(done) == 15
    for g, v in done:
        assert not (g >> 16) & 1
        assert v == P[g], hex(g)
    want = [ln for ln in log if ln.startswith("gen ")][:3]
    assert [ln for ln in res.lines if ln.startswith("gen ")][:3] == want
    # the search's state after three generations is the stored run's fourth line
    fourth = [ln for ln in log if ln.startswith("gen 3 |")][0]
    end = [ln for ln in res.lines if ln.startswith("gen end")][0]
    assert end.split("| evals")[0].replace("gen end", "") == \
        fourth.split("| unique")[0].replace("gen 3", "")
    assert res.lines[-1].startswith("CHAMPION train-30 ")
