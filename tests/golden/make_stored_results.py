"""Writes stored_results.json.gz: the reference's stored per-rule results
that a re-evaluation can reproduce, trimmed to what the checks read.

* nsgp: for each of the 50 instances with an NS-GP baseline run
  (results/baseline/nsgp/<inst>/seed<k>.json, seeds 0-29), every seed's
  final rule `best_tree` (NS-11 terminals) and its stored held-out mean
  reward `test_mean` (as the stored text; 500 scenarios seeded 1,000,000 +
  s, s4_selection/redeploy.cu:14-15).
* traces: for every stored subset-GA trace (results/ga40/results*/
  trace_<inst>_<tag>.json), its instance, extra terminals, mean gain and the
  30 finals (seed, gain, rule) -- what s4_selection/redeploy.cu and
  s4_selection/choice.cu read.

Run from the repository root with the reference at /root/reference:
    python tests/golden/make_stored_results.py
"""
import glob
import gzip
import json
import os

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "stored_results.json.gz")


def main():
    nsgp = {}
    for d in sorted(glob.glob(os.path.join(REF, "results", "baseline", "nsgp", "*"))):
        inst = os.path.basename(d)
        seeds = []
        for s in range(30):
            j = json.load(open(os.path.join(d, "seed%d.json" % s)))
            assert j["instance"] == inst and int(j["seed"]) == s
            tm = j["test_mean"]                # text in most files, a JSON number in some
            seeds.append([j["best_tree"], tm if isinstance(tm, str) else repr(float(tm))])
        nsgp[inst] = seeds
    traces = {}
    for p in sorted(glob.glob(os.path.join(REF, "results", "ga40", "results*", "trace_*.json"))):
        j = json.load(open(p))
        key = os.path.relpath(p, os.path.join(REF, "results", "ga40"))
        traces[key] = {"instance": j["instance"], "vocab_extra": j["vocab_extra"], "mean_gain": j["mean_gain"],
                       "finals": [{"seed": f["seed"], "gain": f["gain"], "rule": f["rule"]} for f in j["finals"]]}
    with gzip.open(OUT, "wt") as f:
        json.dump({"nsgp": nsgp, "traces": traces}, f, separators=(",", ":"), sort_keys=True)
    print("%d NS-GP instances, %d traces -> %s" % (len(nsgp), len(traces), OUT))


if __name__ == "__main__":
    main()
