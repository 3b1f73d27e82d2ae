"""The device scenario sampler (sigp_sample_durations, csrc/sigp_sample.hip)
against its host restatement -- the reference's sampler (sigp/scenarios.py,
pinned by durations.txt and rng.txt) with the correctly rounded log/sincos
of oracle/libcrmath.so injected -- bit for bit; against the reference's own
(libm) sampler it differs only where libm misrounds.  The GP loop fed by
the device sampler equals the loop fed with the restated rows."""
import random

import numpy as np
import pytest

from gpu_util import assert_bitwise
from oracle import CrMath
from sigp import gp as G
from sigp.instance import load_instance, synthetic_clustered, synthetic_mixed
from sigp.rules import EXTRA, FUNCTIONS, NS_TERMINALS, term_mask_from, to_prefix
from sigp.scenarios import gen_durations

pytestmark = pytest.mark.gpu
CR = CrMath()


def _inst(name):
    return {"c101": lambda: load_instance("c101"), "mixed288": lambda: synthetic_mixed(n=288, m=4, seed=1),
            "clustered": lambda: synthetic_clustered(n=100, m=3, seed=1)}[name]()


@pytest.mark.parametrize("inst_name", ["c101", "mixed288"])
@pytest.mark.parametrize("mode,shape", [("gauss", 4.0), ("gamma", 4.0), ("gamma", 2.0), ("gamma", 1.0)])
@pytest.mark.parametrize("seed", [0, 1_000_000, 2 ** 40 + 5])
def test_device_rows_equal_host_restatement(inst_name, mode, shape, seed):
    from sigp.api import FactoryContext, sample_durations
    inst = _inst(inst_name)
    ctx = FactoryContext(inst)
    print("sampler", inst_name, mode, shape, seed, flush=True)
    got = sample_durations(ctx, 23, seed, mode=mode, shape=shape)
    want = gen_durations(inst.durations, 23, seed, mode=mode, shape=shape, math_fns=CR)
    assert_bitwise(got, want, "%s %s rows" % (mode, inst_name))


def test_zero_duration_customers_draw_nothing():
    from sigp.api import FactoryContext, sample_durations
    inst = synthetic_clustered(n=100, m=3, seed=1)
    inst.durations = np.array(inst.durations, dtype=np.float64)
    inst.durations[::7] = 0.0
    ctx = FactoryContext(inst)
    for mode in ("gauss", "gamma"):
        got = sample_durations(ctx, 9, 77, mode=mode)
        want = gen_durations(inst.durations, 9, 77, mode=mode, math_fns=CR)
        assert_bitwise(got, want, mode)
        assert np.all(got[:, ::7] == 0.0)


def test_device_rows_against_the_reference_sampler():
    """The reference's Gaussian sampler uses libm: the device rows agree
    with it except where libm misrounds log or sincos, by a few ulps."""
    from sigp.api import FactoryContext, sample_durations
    inst = load_instance("c101")
    ctx = FactoryContext(inst)
    got = sample_durations(ctx, 200, 1_000_000)
    ref = gen_durations(inst.durations, 200, 1_000_000)
    same = float(np.mean(got == ref))
    print("identical to the libm sampler: %.4f of %d values" % (same, got.size), flush=True)
    assert same > 0.98
    d = got != ref
    assert np.all(np.abs(got[d] - ref[d]) <= 4 * np.spacing(np.abs(ref[d])))


def test_gp_loop_with_device_sampler_equals_restated_rows():
    """Two runs x 128 rules x 4 Gamma scenarios x 4 generations, rows drawn
    inside the loop from Mt19937(run_seed + g + t) (the reference's
    schedule), in two calls (gen_offset), graph-captured."""
    from sigp.api import FactoryContext
    from sigp.gp_device import DeviceGP
    inst = synthetic_clustered(n=100, m=3, seed=1)
    ctx = FactoryContext(inst)
    terms = list(NS_TERMINALS) + list(EXTRA)
    mask = term_mask_from(terms)
    tc, ec = G.TreeConfig(terms, list(FUNCTIONS)), G.EvolutionConfig(population_size=128)
    seeds, T, gens = [3, 9], 4, 4
    run_seed = [s * 10000 for s in seeds]
    rows = np.concatenate([gen_durations(inst.durations, T, run_seed[r] + g, mode="gamma", math_fns=CR)
                           for g in range(gens) for r in range(len(seeds))])
    host = DeviceGP(ctx, tc, ec, seeds=seeds, T=T, term_mask=mask)
    host.init()
    want = host.run(rows, generations=gens)
    dev = DeviceGP(ctx, tc, ec, seeds=seeds, T=T, term_mask=mask)
    dev.init()
    dev.set_sampler(run_seed, mode="gamma", gen_offset=0)
    a = dev.run(None, generations=2, step_last=True)
    dev.set_sampler(run_seed, mode="gamma", gen_offset=2)
    b = dev.run(None, generations=2)
    assert_bitwise(np.concatenate([a.fitness, b.fitness]), want.fitness, "device-sampled loop")
    assert [[to_prefix(t) for t in p] for p in dev.population()] == \
        [[to_prefix(t) for t in p] for p in host.population()]
