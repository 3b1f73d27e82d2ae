"""The GA terminal-subset search (sigp/ga_search.py, restating
s4_selection/ga.cu:267-520) replays the reference's stored search runs line
for line (CPU only).

For each instance whose evaluation cache the reference kept
(tests/golden/gasearch/<inst>.evcache.gz: every proxy and confirmation value
its search computed) the search is run with the proxy answered from that
cache.  It starts from the entries the run had loaded when it began (its log
says how many: the run resumed after an interruption), so the evaluation
counts it prints are the tool's.  Every generation line (best genome and its
mean, population mean, distinct genomes, evaluations), every reinjection
and the stagnation stop, the final line, the confirmed finalists and the
champion must equal the stored log (tests/golden/gasearch/<inst>_search.log.gz).
"""
import gzip
import os
import re

import pytest

from sigp.ga_search import ga_search, read_cache

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gasearch")
INSTANCES = ["c204", "c205", "pr14", "pr15", "rc202", "rc205"]
KEEP = re.compile(r"^(gen \d+ \||gen end \||  reinject|stop: |  \d+\.\d\d \(proxy|CHAMPION )")


def _stored(inst):
    with gzip.open(os.path.join(GOLDEN, inst + "_search.log.gz"), "rt") as f:
        log = [ln.rstrip("\n") for ln in f]
    with gzip.open(os.path.join(GOLDEN, inst + ".evcache.gz"), "rt") as f:
        header, entries = read_cache(f)
    return log, header, entries


def _options(log, header):
    universe = re.search(r"uni=(\S+)", header).group(1).split(",")
    head = next(ln for ln in log if ln.startswith("# ga "))
    m = re.search(r"seeds (\w+) \(\d+\)( nofull)? \| tree-13 uniform-x pm ([\d.]+) \| gens (\d+) stag (\d+)", head)
    loaded = 0
    for ln in log:
        mm = re.match(r"# cache \S+ \| (\d+) entries loaded", ln)
        if mm:
            loaded = int(mm.group(1))
    top = int(re.search(r"confirming top-(\d+) at", "\n".join(log)).group(1))
    return dict(universe=universe, seed_mode=m.group(1), nofull=bool(m.group(2)), pm=float(m.group(3)),
                gens=int(m.group(4)), stag_k=int(m.group(5)), top_n=top), loaded


@pytest.mark.parametrize("inst", INSTANCES)
def test_search_replays_stored_log(inst):
    log, header, entries = _stored(inst)
    opts, loaded = _options(log, header)
    P = {g: v for k, g, v in entries if k == "P"}
    C = {g: v for k, g, v in entries if k == "C"}
    cache = {g: v for k, g, v in entries[:loaded] if k == "P"}
    cache_conf = {g: v for k, g, v in entries[:loaded] if k == "C"}
    asked = []

    def proxy(gs):
        asked.extend(gs)
        return [P[g] for g in gs]

    res = ga_search(opts.pop("universe"), proxy, lambda gs: [C[g] for g in gs], cache=cache,
                    cache_conf=cache_conf, **opts)
    want = [ln for ln in log if KEEP.match(ln)]
    got = [ln for ln in res.lines if KEEP.match(ln)]
    assert got == want
    # the run asked exactly for the proxy entries the stored cache appended after loading
    assert sorted(asked) == sorted(g for k, g, _ in entries[loaded:] if k == "P")


def test_search_is_deterministic_and_uses_cache():
    """Same proxy, same search; a second run over the first one's cache
    evaluates nothing and prints the same generations."""
    log, header, entries = _stored("c205")
    opts, _ = _options(log, header)
    P = {g: v for k, g, v in entries if k == "P"}
    C = {g: v for k, g, v in entries if k == "C"}
    universe = opts.pop("universe")
    a = ga_search(universe, lambda gs: [P[g] for g in gs], lambda gs: [C[g] for g in gs], **opts)
    b = ga_search(universe, lambda gs: 1 / 0, lambda gs: 1 / 0, cache=a.cache, cache_conf=a.cache_conf, **opts)
    gen = lambda r: [re.sub(r"evals \d+", "", ln) for ln in r.lines if ln.startswith("gen")]
    assert gen(a) == gen(b) and b.evals == 0 and a.champion == b.champion
