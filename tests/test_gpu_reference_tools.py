"""The reference's own GPU tools, compiled unchanged against
include/sigp_factory.hpp (oracle/Makefile, target tools), reproduce the
reference's stored results on MI355X:

* s3_generation/gen_trace.cu: the 30-seed GP trace of an instance (every
  seed's final rule and test gain, every generation's best training fitness,
  size and extra terminals) equals results/ga40/results/trace_<inst>_Sga.json
  field for field;
* include/sigp_gp.hpp: evolve_one_device (the GP loop on the device) returns
  the population of the reference's own generation loop (its engine/gp.hpp
  operators on the host) tree for tree;
* s4_selection/ga.cu: the tool's own --gate passes; the proxy fitness of a
  reduced initial pool (base + 3 singles x 16 seeds x 51 generations)
  equals the stored cache lines of results/ga40/logs/c204.evcache; the
  whole initial pool (19 terminal subsets, about 3.5 minutes, with
  SIGP_SLOW_TESTS=1) equals it except the subset {ATRISK}, whose stored
  values come from the reference's CUDA build (its erf differs from glibc,
  verify/verify_cuda.cu:145-149; this library is bitwise to the CPU
  simulator).

The binaries exist where the reference sources were present at build time;
elsewhere these tests skip.  Fixtures: tests/golden/traces, gacache and
cap2-results.csv (the reference's baseline table the tools read).
"""
import gzip
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "bin")
GOLDEN = os.path.join(ROOT, "tests", "golden")
UNIVERSE = ("REGRET,NSROB,MAXN,REACH2,NSTW,FRAGCNT,FRAGSCORE,PBUST,REACH1,REACH3,REACH5,SAT,VFRAC,NFEAS,"
            "HARVEST,BOLSAO,ATRISK,REACH2ROB")


def _env():
    env = dict(os.environ)
    env["GITC_DATA"] = os.path.join(GOLDEN, "instances")
    env["GITC_CAP2"] = os.path.join(GOLDEN, "cap2-results.csv")
    return env


def _tool(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.skip("%s not built (needs the reference sources at build time)" % name)
    return path


@pytest.mark.parametrize("inst", ["rc101", "pr12"])
def test_gen_trace_reproduces_stored_trace(inst, tmp_path):
    want = json.load(gzip.open(os.path.join(GOLDEN, "traces", "trace_%s_Sga.json.gz" % inst), "rt"))
    terms = ",".join(want["vocab_extra"])
    print("gen_trace", inst, "--terms", terms, flush=True)
    res = subprocess.run([_tool("gen_trace"), inst, "--terms", terms, "--tag", "Sga", "--seeds", "30"],
                         cwd=tmp_path, env=_env(), capture_output=True, text=True, timeout=300)
    print(res.stdout[-400:], flush=True)
    assert res.returncode == 0, res.stderr[-2000:]
    got = json.load(open(tmp_path / "results" / ("trace_%s_Sga.json" % inst)))
    for key in ("instance", "vocab_extra", "mean_gain", "std", "seeds_positive", "best_seed_gain"):
        assert got[key] == want[key], key
    for a, b in zip(got["finals"], want["finals"]):
        assert a == b, "seed %d final" % b["seed"]
    for s, (a, b) in enumerate(zip(got["trace"], want["trace"])):
        assert a == b, "seed %d trace" % s
    assert len(got["finals"]) == len(want["finals"]) == 30


def _cache(lines):
    P = {}
    for ln in lines:
        p = ln.split()
        if p and p[0] == "P":
            P[int(p[1], 16)] = p[3:]
    return P


def _run_ga(args, cwd, timeout):
    """Runs the reference's ga with its progress lines echoed as they come
    (line-buffered, s4_selection/ga.cu:52,202); returns (rc, output)."""
    proc = subprocess.Popen([_tool("ga")] + args, cwd=cwd, env=_env(), stdout=subprocess.PIPE,
                            stderr=subprocess.STDOUT, text=True)
    out = []
    for line in proc.stdout:
        print(line.rstrip()[:200], flush=True)
        out.append(line)
    return proc.wait(timeout=timeout), "".join(out)


def test_ga_gate_passes(tmp_path):
    """The tool's own gate (s4_selection/ga.cu:267-290): its batched
    evaluate_grouped_masks path against per-genome evaluate_grouped /
    evaluate, 5 genomes (base, all 18, two singles, a triple) x 4 seeds x 5
    generations, on this library; every value must match."""
    rc, out = _run_ga(["c204", "--universe", UNIVERSE, "--gate"], tmp_path, 600)
    assert rc == 0, out[-2000:]
    assert "GATE_PASS" in out and "mismatches 0" in out


def test_ga_reduced_initial_pool_reproduces_stored_cache(tmp_path):
    """The GA's initial pool over a 3-terminal universe (REGRET, NSROB,
    MAXN): the base genome and the three singles, each 16 seeds x 51
    generations of GP on the library.  A genome's vocabulary is NS-11 plus
    its terminals in universe order, so these four are the stored run's
    genomes 0, 1, 2 and 4, whose cache lines (16 hexfloats each) must match."""
    rc, out = _run_ga(["c204", "--universe", "REGRET,NSROB,MAXN", "--seeds", "neutral", "--nofull", "--pm", "0.5",
                       "--gens", "0", "--stag", "10", "--top", "1", "--confirm", "1", "--bmax", "4",
                       "--cache", "c204.evcache"], tmp_path, 600)
    assert rc == 0, out[-2000:]
    got = _cache(open(tmp_path / "c204.evcache"))
    want = _cache(gzip.open(os.path.join(GOLDEN, "gacache", "c204.evcache.gz"), "rt"))
    for g in (0, 1, 2, 4):
        assert g in got and g in want
        assert got[g] == want[g], hex(g)


@pytest.mark.skipif(os.environ.get("SIGP_SLOW_TESTS") != "1", reason="about 3.5 minutes; set SIGP_SLOW_TESTS=1")
def test_ga_initial_pool_reproduces_stored_cache(tmp_path):
    # The tool runs for minutes and prints a line per batch of genomes
    # (line-buffered, s4_selection/ga.cu:52,202), echoed here as they come.
    # Batches of at most 4 genomes instead of the stored run's 13 (a line
    # about every minute); batching does not change a genome's values, which
    # the tool's own --gate compares.
    proc = subprocess.Popen([_tool("ga"), "c204", "--universe", UNIVERSE, "--seeds", "neutral", "--nofull",
                             "--pm", "0.5", "--gens", "0", "--stag", "10", "--top", "1", "--confirm", "1",
                             "--bmax", "4", "--cache", "c204.evcache"],
                            cwd=tmp_path, env=_env(), stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    tail = []
    for line in proc.stdout:
        print(line.rstrip()[:200], flush=True)
        tail = (tail + [line])[-40:]
    assert proc.wait(timeout=1000) == 0, "".join(tail)
    got = _cache(open(tmp_path / "c204.evcache"))
    want = _cache(gzip.open(os.path.join(GOLDEN, "gacache", "c204.evcache.gz"), "rt"))
    initial = [0] + [1 << b for b in range(18)]
    atrisk = 1 << 16
    for g in initial:
        assert g in got and g in want
        if g != atrisk:
            assert got[g] == want[g], hex(g)


@pytest.mark.parametrize("args", [["c101"], ["rc101", "--terms", "REACH2,FRAGCNT", "--seed", "3", "--samples", "5",
                                              "--gens", "20"]],
                         ids=["c101-ns11-51gens", "rc101-ext-20gens"])
def test_evolve_one_device_equals_the_reference_loop(args, tmp_path):
    """include/sigp_gp.hpp's evolve_one_device against the reference's own
    generation loop (engine/gp.hpp operators on the host, FactoryContext
    evaluation): identical final populations (oracle/gp_swap_check.cpp)."""
    path = os.path.join(ROOT, "oracle", "_ref", "gp_swap_check")
    if not os.path.exists(path):
        pytest.skip("gp_swap_check not built (needs the reference sources at build time)")
    res = subprocess.run([path] + args, cwd=tmp_path, env=_env(), capture_output=True, text=True, timeout=300)
    print(res.stdout, flush=True)
    assert res.returncode == 0, res.stdout + res.stderr[-2000:]
