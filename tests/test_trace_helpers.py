"""The device trace's host helpers (sigp/trace_device.py) against the stored
traces' own fields (tests/golden/traces): every final rule's extra terminals
(`ex`) and size (`sz`), re-derived from its prefix string, equal the stored
ones, and the stored gains and means are fixed points of the reference's
rounding (engine/jwrite.hpp:10-20, restated as round_half_even)."""
import gzip
import json
import os

import pytest

from sigp.rules import NS_TERMINALS, parse_prefix, to_prefix, tree_size
from sigp.trace_device import extras_of, round_half_even

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("inst", ["rc101", "pr12", "c202"])
def test_extras_size_and_prefix_round_trip(inst):
    t = json.load(gzip.open(os.path.join(GOLDEN, "traces", "trace_%s_Sga.json.gz" % inst), "rt"))
    base = set(NS_TERMINALS)
    for f in t["finals"]:
        rule = parse_prefix(f["rule"])
        assert to_prefix(rule) == f["rule"]
        assert tree_size(rule) == f["sz"]
        assert extras_of(rule, base) == f["ex"]
    for key in ("mean_gain", "std", "best_seed_gain"):
        assert round_half_even(t[key], 2) == t[key]
    for f in t["finals"]:
        assert round_half_even(f["gain"], 2) == f["gain"]


def test_round_half_even_ties_and_signs():
    assert round_half_even(0.125, 2) == 0.12      # exact tie (0.125 is a double): to even
    assert round_half_even(0.375, 2) == 0.38
    assert round_half_even(-0.125, 2) == -0.12
    assert round_half_even(2.5, 0) == 2.0 and round_half_even(3.5, 0) == 4.0
    assert round_half_even(1109.25, 1) == 1109.2
    assert round_half_even(7.0, 1) == 7.0
