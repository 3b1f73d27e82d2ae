"""The Python mirror of the reference's host code against its fixtures:
instance readers (parser.txt), rule compiler (compile.txt), random streams
(rng.txt), duration sampling (durations.txt / sim.txt rows) and the GP
operators (gp.txt populations)."""
import gzip
import os
import random

import numpy as np
import pytest

from golden_data import GOLDEN, bits, compile_cases, d2h, gp_cases, h2d, sim_cases
from sigp import gp as G
from sigp.instance import load_instance
from sigp.rules import (EXTRA, FUNCS_RMOD, FUNCTIONS, NS_TERMINALS, compile_tree, pack_trees,
                        parse_prefix, stack_depth, to_prefix, term_mask_from)
from sigp.scenarios import gen_durations


def test_parser_fixture():
    n_inst = fields = 0
    inst = None
    k = 0
    with gzip.open(os.path.join(GOLDEN, "parser.txt.gz"), "rt") as fh:
        for ln in fh:
            parts = ln.split()
            if not parts:
                continue
            if parts[0] == "#":
                inst = load_instance(parts[2])
                k = 0
                n_inst += 1
                continue
            fields += 1
            if parts[0] == "name":
                assert inst.name == parts[1]
            elif parts[0] == "n":
                assert len(inst.pois) == int(parts[1])
            elif parts[0] == "tmax":
                assert d2h(inst.tmax) == parts[1]
            elif parts[0] == "poi":
                p = inst.pois[k]
                k += 1
                assert p.id == int(parts[1])
                assert [d2h(v) for v in (p.x, p.y, p.score, p.open_time, p.close_time, p.duration)] == parts[2:8]
    assert n_inst == 50 and fields > 5000


def test_dist_matrix_is_sqrt_form():
    inst = load_instance("pr01")
    cx = np.array([p.x for p in inst.pois]); cy = np.array([p.y for p in inst.pois])
    a, b = 3, 17
    assert inst.dist[a, b] == np.sqrt((cx[a] - cx[b]) ** 2 + (cy[a] - cy[b]) ** 2)
    assert np.array_equal(bits(inst.dist), bits(inst.dist.T))


@pytest.mark.parametrize("case", compile_cases(), ids=lambda c: str(len(c["tree"])))
def test_compile_fixture(case):
    t = parse_prefix(case["tree"])
    assert to_prefix(t) == case["tree"]
    ops, args = compile_tree(t)
    assert ops == case["ops"]
    assert [x for pr in args for x in pr] == case["args"]


def test_compile_constants_and_pack():
    ops, args = compile_tree(parse_prefix("(rmod 0.0005 (- -1.2345 TA))"))
    assert ops == [-2, -2, -1, 1, 6]
    assert args[0] == (1, 0) and args[1] == (-1235, 0) and args[2] == (4, 0)   # llround half away
    pk = pack_trees([parse_prefix("SCORE"), parse_prefix("(+ TA TR)")])
    assert list(pk.offsets) == [0, 1] and list(pk.lengths) == [1, 3]
    assert list(pk.args) == [0, 0, 4, 0, 5, 0, 0, 1]
    assert stack_depth([-1, -1, -1, 0, 0]) == 3


def _rng_lines():
    with open(os.path.join(GOLDEN, "rng.txt")) as fh:
        for ln in fh:
            if ln.strip():
                yield ln.split()


def test_rng_fixture_is_cpython_random():
    """engine/rng.hpp restates CPython's random.Random: rng.txt pins it."""
    rng = None
    n = 0
    for parts in _rng_lines():
        tag = parts[0]
        if tag == "#":
            rng = random.Random(int(parts[2]))
            continue
        n += 1
        if tag in ("random", "mix_random"):
            got = d2h(rng.random())
            assert got == parts[1]
        elif tag.startswith("getrandbits_"):
            assert str(rng.getrandbits(int(tag[12:]))) == parts[1]
        elif tag in ("randint_2_8", "mix_randint"):
            assert str(rng.randint(2, 8)) == parts[1]
        elif tag == "choice29":
            assert str(rng._randbelow(29)) == parts[1]
        elif tag == "mix_choice":
            assert str(rng._randbelow(61)) == parts[1]
        elif tag in ("sample1024_7", "sample50_7"):
            s = rng.sample(range(1024 if tag == "sample1024_7" else 50), 7)
            assert [str(v) for v in s] == parts[1:]
        elif tag == "gauss":
            assert d2h(rng.gauss(10.0, 2.0)) == parts[1]
        elif tag == "mix_gauss":
            assert d2h(rng.gauss(5.0, 1.5)) == parts[1]
        else:
            raise AssertionError("unknown tag " + tag)
    assert n == 6416


def test_durations_fixture_exact_and_sim_rows_within_one_ulp():
    """durations.txt (3 instances x 3 base seeds x 5 scenarios) is reproduced
    bit for bit; the sampled rows inside sim.txt were written on another libm
    and agree to within 1 ulp (the simulator gates feed those rows directly)."""
    with open(os.path.join(GOLDEN, "durations.txt")) as fh:
        inst = base = None
        s = n = 0
        for ln in fh:
            parts = ln.split()
            if not parts:
                continue
            if parts[0] == "#":
                inst = load_instance(parts[2]); base = int(parts[4]); s = 0
            elif parts[0] == "dur":
                want = np.array([h2d(v) for v in parts[1:]])
                got = gen_durations(inst.durations, 1, base + s)[0]
                assert np.array_equal(bits(got), bits(want))
                s += 1
                n += want.size
    assert n == 3720
    for c in sim_cases()[::3]:
        inst = load_instance(c["inst"])
        got = gen_durations(inst.durations, 20, 1_000_000)
        diff = np.abs(bits(got).astype(np.int64) - bits(c["sampled"]).astype(np.int64))
        assert diff.max() <= 1


def test_durations_match_reference_sampler(reference):
    """Against the reference's own gen_durations (oracle/_ref): pr01 has
    per-customer durations, so every draw is a distinct gauss call."""
    for name in ("pr01", "c101"):
        inst = load_instance(name)
        for base in (0, 1_000_000, 12345):
            got = gen_durations(inst.durations, 40, base)
            want = reference.gen_durations(inst.durations, 40, base)
            assert np.array_equal(bits(got), bits(want))


def test_gp_populations_fixture():
    """gp.txt: ramped_half_and_half + 2 gp_step generations from Mt19937(0),
    with the fixture's fitness values, reproduce all 6,144 prefix strings."""
    for case in gp_cases():
        terms = list(NS_TERMINALS) + (list(EXTRA) if case["kind"] == "ext29rmod" else [])
        funcs = FUNCS_RMOD if case["kind"] == "ext29rmod" else FUNCTIONS
        tc = G.TreeConfig(terms, list(funcs))
        cfg = G.EvolutionConfig()
        rng = random.Random(0)
        pop = G.ramped_half_and_half(tc, rng, cfg.population_size)
        for gi, g in enumerate(case["gens"]):
            assert [to_prefix(t) for t in pop] == g["pop"], "generation %d" % gi
            if gi < 2:
                pop = G.gp_step(pop, list(g["fits"]), rng, tc, cfg)


def test_term_mask():
    m = term_mask_from(NS_TERMINALS + EXTRA + ["ERC", "WLOSS"])
    assert m.sum() == 29 and m[11] == 1
    with pytest.raises(ValueError):
        term_mask_from(["NOPE"])
