"""The library against the reference's stored per-rule results on every
published instance (tests/golden/stored_results.json.gz):

* NS-GP baseline: all 40 instances x 30 seeds' final rules
  (results/baseline/nsgp/<inst>/seed<k>.json), re-scored over the 500
  held-out scenarios seeded 1,000,000 + s with the NS-11 mask, give the
  stored `test_mean` to its stored digits;
* subset-GA traces: all 51 stored traces x 30 finals
  (results/ga40/results*/trace_*.json), re-scored the way
  s4_selection/redeploy.cu:29-46 re-scores a trace's best, give the stored
  gain 100 (mean - km) / km to its two decimals (km: the NS-GP mean of
  cap2-results.csv);
* the reference's own `redeploy` tool, compiled against this library,
  prints SEALED for every trace, and its `choice` tool prints the training
  means the oracle gives (oracle/_ref/bin; skipped where not built).
"""
import csv
import gzip
import json
import os
import re
import subprocess

import numpy as np
import pytest

from sigp.instance import load_instance
from sigp.rules import NS_TERMINALS, pack_trees, parse_prefix, term_mask_from
from sigp.scenarios import gen_durations
from test_stored_results import N_TEST, TEST_SEED, stored, within_stored_digits

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
BIN = os.path.join(ROOT, "oracle", "_ref", "bin")
THREADS = min(16, os.cpu_count() or 1)


def _km():
    with open(os.path.join(GOLDEN, "cap2-results.csv")) as f:
        return {r["instance"]: float(r["baseline_nsgp_mean30"]) for r in csv.DictReader(f)}


def _env():
    env = dict(os.environ)
    env["GITC_DATA"] = os.path.join(GOLDEN, "instances")
    env["GITC_CAP2"] = os.path.join(GOLDEN, "cap2-results.csv")
    return env


def test_library_reproduces_every_stored_nsgp_test_mean():
    from sigp.api import FactoryContext
    bad, n = [], 0
    for name, seeds in sorted(stored()["nsgp"].items()):
        inst = load_instance(name)
        ctx = FactoryContext(inst)
        pk = pack_trees([parse_prefix(t) for t, _ in seeds])
        test = gen_durations(inst.durations, N_TEST, TEST_SEED)
        _, mean = ctx.evaluate_means(pk, test, N_TEST, term_mask_from(NS_TERMINALS))
        bad += [(name, k, float(g), w) for k, (g, (_, w)) in enumerate(zip(mean, seeds))
                if not within_stored_digits(g, w)]
        n += len(seeds)
        ctx.close()
    print("NS-GP stored test means: %d rules, %d differ" % (n, len(bad)), flush=True)
    assert n == 1200 and not bad, bad[:5]


def test_library_reproduces_every_stored_trace_gain():
    from sigp.api import FactoryContext
    km = _km()
    bad, n = [], 0
    for key, t in sorted(stored()["traces"].items()):
        name = t["instance"]
        inst = load_instance(name)
        ctx = FactoryContext(inst)
        finals = t["finals"]
        pk = pack_trees([parse_prefix(f["rule"]) for f in finals])
        test = gen_durations(inst.durations, N_TEST, TEST_SEED)
        _, mean = ctx.evaluate_means(pk, test, N_TEST, term_mask_from(list(NS_TERMINALS) + t["vocab_extra"]))
        for f, m in zip(finals, mean):
            gain = 100.0 * (m - km[name]) / km[name]
            if abs(gain - f["gain"]) > 0.005 + 1e-9:
                bad.append((key, f["seed"], round(gain, 4), f["gain"]))
        n += len(finals)
        ctx.close()
    print("trace finals: %d rules, %d differ" % (n, len(bad)), flush=True)
    assert n == 1530 and not bad, bad[:5]


def _tool(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.skip("%s not built (needs the reference sources at build time)" % name)
    return path


def _write_trace(tmp_path, key, t):
    path = tmp_path / os.path.basename(key)
    path.write_text(json.dumps(t))
    return str(path)


def test_redeploy_seals_every_stored_trace(tmp_path):
    tool = _tool("redeploy")
    sealed = 0
    for key, t in sorted(stored()["traces"].items()):
        sub = tmp_path / key.split("/")[0]
        sub.mkdir(exist_ok=True)
        path = _write_trace(sub, key, t)
        res = subprocess.run([tool, t["instance"], path], cwd=tmp_path, env=_env(), capture_output=True,
                             text=True, timeout=120)
        assert res.returncode == 0, res.stderr[-1000:]
        line = res.stdout.strip().splitlines()[-1]
        assert line.endswith("SEALED"), line
        sealed += 1
    print("redeploy: %d traces sealed" % sealed, flush=True)


@pytest.mark.parametrize("name", ["c101", "r101"])
def test_choice_prints_the_oracle_training_means(oracle, name, tmp_path):
    """choice <inst> <Sga trace> <FULL18 trace> --backup: the training mean
    of each column (30 rules x 64 scenarios seeded 2,000,000 + s,
    s4_selection/choice.cu:15-16,43-53) equals the oracle's, and the chosen
    column is the one with the largest."""
    tool = _tool("choice")
    S = stored()
    inst = load_instance(name)
    nsdir = tmp_path / "results" / "baseline" / "nsgp" / name
    nsdir.mkdir(parents=True)
    for k, (tree, _) in enumerate(S["nsgp"][name]):
        (nsdir / ("seed%d.json" % k)).write_text(json.dumps({"best_tree": tree}))
    keys = ["results/trace_%s_Sga.json" % name, "results_full18/trace_%s_FULL18.json" % name]
    paths = [_write_trace(tmp_path, k, S["traces"][k]) for k in keys]
    res = subprocess.run([tool, name] + paths + ["--backup"], cwd=tmp_path, env=_env(), capture_output=True,
                         text=True, timeout=120)
    print(res.stdout, flush=True)
    assert res.returncode == 0, res.stderr[-1000:]
    got = [float(x) for x in re.findall(r"mean-TRAIN\s+(-?[\d.]+)", res.stdout)]

    train = gen_durations(inst.durations, 64, 2000000)
    zp = oracle.zstats(inst)
    cols = [([parse_prefix(t) for t, _ in S["nsgp"][name]], list(NS_TERMINALS))]
    for k in keys:
        t = S["traces"][k]
        finals = sorted(t["finals"], key=lambda f: f["seed"])
        cols.append(([parse_prefix(f["rule"]) for f in finals], list(NS_TERMINALS) + t["vocab_extra"]))
    want = []
    for rules, vocab in cols:
        fit = oracle.evaluate(inst, pack_trees(rules), train, 64, term_mask_from(vocab), zp,
                              threads=THREADS).reshape(len(rules), 64)
        acc = 0.0
        for p in range(len(rules)):
            m = 0.0
            for s in range(64):
                m += fit[p, s]
            acc += m / 64
        want.append(acc / len(rules))
    assert got == [float("%.2f" % w) for w in want], (got, want)
    labels = ["NS-GP(backup)"] + [os.path.splitext(os.path.basename(k))[0] for k in keys]
    best = labels[int(np.argmax(want))]
    assert "CROSS-VALIDATORY CHOICE (train-side): %s " % best in res.stdout
