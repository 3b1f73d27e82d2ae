"""The committed work models (tests/golden/work_*.json) still describe their
benchmark workloads: same arrays (digest), and the counting oracle gives the
same operation counts on a subset of rules (of (instance, rule) pairs for
the multi-instance batch)."""
import json
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_work_model as mwm  # noqa: E402

KINDS = ["plain", "large", "batched"]


@pytest.fixture(scope="module", params=KINDS)
def case(request):
    kind = request.param
    with open(mwm.path(kind)) as f:
        model = json.load(f)
    return kind, model, mwm.workload(kind)


def test_workload_digest(case):
    from sigp.workloads import workload_digest
    kind, model, wl = case
    assert model["digest"] == workload_digest(wl)
    n_inst = len(wl.insts) if kind == "batched" else 1
    assert len(model["ops"]) == n_inst * wl.packed.n_rules
    n = len(model["ops"]) * model["scenario_sample"]["count"]
    assert model["ops_per_sim"] == pytest.approx(sum(model["ops"]) / n, rel=1e-12)
    if kind == "plain":
        assert len(model["ops_full"]) == wl.packed.n_rules


def test_recount_subset(case):
    from oracle import Oracle
    kind, model, wl = case
    O = Oracle(counting=True)
    P = wl.packed.n_rules
    if kind == "batched":
        picks = [(k, p) for k, p in [(0, 0), (7, 100), (19, 255), (33, 31), (39, 128)]]
    else:
        picks = [(0, p) for p in range(0, P, P // 8)][:8 if kind == "plain" else 4]
    for k, p in picks:
        inst = wl.insts[k] if kind == "batched" else wl.inst
        zp = O.zstats(inst)
        S = mwm.sample_rows(wl, kind, k)
        assert mwm.count_rule(O, wl, p, S, zp, True, inst) == model["ops"][k * P + p]
        if kind == "plain":
            assert mwm.count_rule(O, wl, p, S, zp, False) == model["ops_full"][p]
    if kind == "plain":
        # pruning never adds work
        assert all(a <= b for a, b in zip(model["ops"], model["ops_full"]))
