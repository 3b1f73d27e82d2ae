"""The reference's GP trace tool (s3_generation/gen_trace.cu) restated with
the GP loop on the device (sigp/trace_device.py): the 30-seed trace of an
instance -- every seed's final rule and held-out gain, every generation's
best training fitness, size and extra terminals, the summary -- equals the
stored results/ga40/results/trace_<inst>_Sga.json field for field."""
import csv
import gzip
import json
import os
import time

import pytest

from sigp.instance import load_instance

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("inst", ["rc101", "pr12", "c202"])
def test_device_trace_equals_the_stored_trace(inst):
    from sigp.api import FactoryContext
    from sigp.trace_device import trace
    want = json.load(gzip.open(os.path.join(GOLDEN, "traces", "trace_%s_Sga.json.gz" % inst), "rt"))
    with open(os.path.join(GOLDEN, "cap2-results.csv")) as f:
        km = {r["instance"]: float(r["baseline_nsgp_mean30"]) for r in csv.DictReader(f)}[inst]
    ctx = FactoryContext(load_instance(inst))
    t0 = time.perf_counter()
    got = trace(ctx, inst, want["vocab_extra"], km)
    print("device trace %s: 30 seeds x 51 generations in %.2f s" % (inst, time.perf_counter() - t0), flush=True)
    for key in ("instance", "vocab_extra", "mean_gain", "std", "seeds_positive", "best_seed_gain"):
        assert got[key] == want[key], key
    assert len(got["finals"]) == len(want["finals"]) == 30
    for a, b in zip(got["finals"], want["finals"]):
        assert a == b, "seed %d final" % b["seed"]
    for s, (a, b) in enumerate(zip(got["trace"], want["trace"])):
        assert a == b, "seed %d trace" % s


def _stored_traces():
    from test_stored_results import stored
    return stored()["traces"]


# Traces whose vocabulary includes ATRISK are left out: the stored traces come
# from the reference's CUDA build, whose erf differs from glibc's in the last
# place (verify/verify_cuda.cu:145-149), and over 51 generations such a
# difference can change the populations (9 of those 16 traces differ; this
# library is bitwise to the reference's CPU simulator on ATRISK, sim.txt).
ALL = sorted(k for k, t in _stored_traces().items() if "ATRISK" not in t["vocab_extra"])
SOME = ["results/trace_c105_Sga.json", "results/trace_r203_Sga.json", "results/trace_pr14_Sga.json"]


@pytest.mark.parametrize("key", [k for k in ALL if os.environ.get("SIGP_SLOW_TESTS") == "1" or k in SOME])
def test_device_trace_finals_equal_every_stored_trace(key):
    """Every stored trace without ATRISK (35 of the 51;
    tests/golden/stored_results.json.gz keeps their finals: seed, gain,
    rule, and the mean gain): the device trace gives the same final rules and
    gains.  Three by default; all 35 with SIGP_SLOW_TESTS=1 (about 2
    minutes)."""
    from sigp.api import FactoryContext
    from sigp.trace_device import trace
    want = _stored_traces()[key]
    inst = want["instance"]
    with open(os.path.join(GOLDEN, "cap2-results.csv")) as f:
        km = {r["instance"]: float(r["baseline_nsgp_mean30"]) for r in csv.DictReader(f)}[inst]
    t0 = time.perf_counter()
    got = trace(FactoryContext(load_instance(inst)), inst, want["vocab_extra"], km)
    print("device trace %s: %.2f s" % (key, time.perf_counter() - t0), flush=True)
    assert got["mean_gain"] == want["mean_gain"]
    assert [(f["seed"], f["gain"], f["rule"]) for f in got["finals"]] == \
        [(f["seed"], f["gain"], f["rule"]) for f in want["finals"]]
