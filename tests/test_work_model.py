"""The committed work model (tests/golden/work_plain_small.json) still
describes the benchmark workload: same arrays, and the counting oracle gives
the same per-rule operation counts on a subset of rules."""
import json
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_work_model as mwm  # noqa: E402


@pytest.fixture(scope="module")
def model():
    with open(mwm.OUT) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def wl():
    from sigp.workloads import plain_small
    return plain_small()


def test_workload_digest(model, wl):
    assert model["digest"] == mwm.workload_digest(wl)
    assert len(model["ops"]) == wl.packed.n_rules == len(model["ops_full"])
    n = wl.packed.n_rules * model["scenario_sample"]["count"]
    assert model["ops_per_sim"] == pytest.approx(sum(model["ops"]) / n, rel=1e-12)


def test_recount_subset(model, wl):
    from oracle import Oracle
    O = Oracle(counting=True)
    zp = O.zstats(wl.inst)
    S = mwm.sample_rows(wl)
    for p in range(0, wl.packed.n_rules, 128):
        assert mwm.count_rule(O, wl, p, S, zp, True) == model["ops"][p]
        assert mwm.count_rule(O, wl, p, S, zp, False) == model["ops_full"][p]
    # pruning never adds work
    assert all(a <= b for a, b in zip(model["ops"], model["ops_full"]))
