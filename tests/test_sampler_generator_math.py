"""Mathematical pins for the two components the reference does not have
(CPU only): the Gamma service-time sampler and the synthetic instance
generators.  The device sampler is bit-identical to the host sampler tested
here (tests/test_gpu_sampler.py), so these pins carry over to it.

Gamma(k, theta) by Marsaglia & Tsang (sigp/scenarios.py::gamma_draw) is
checked against what the distribution fixes, not against a second copy of
the method:
  * sample mean and variance against k*theta and k*theta^2, within 5
    standard errors (the variance's from the Gamma fourth moment);
  * Kolmogorov-Smirnov distance to the Gamma CDF (scipy.stats.gamma) below
    the 0.1 % critical value;
  * for integer k, the same distance to the closed-form Erlang CDF
    1 - sum_{i<k} e^{-x/theta} (x/theta)^i / i!, and a two-sample KS test
    against sums of k exponentials -theta*log(U) drawn from an independent
    stream;
  * shape < 1 (the Gamma(k+1) * U^(1/k) boost) the same way.
Tried on the host with planted mistakes: d = k - 1/2, a lost rejection step,
no squeeze test and a squeeze constant ten times too small (which accepts
points the exact test would reject) each fail at least one of these tests.
A larger squeeze constant is not a mistake: it passes fewer points early and
the exact test decides the rest, so the law is unchanged.
"""
import math
import random

import numpy as np
import pytest
from scipy import stats

from sigp.instance import synthetic_clustered, synthetic_mixed, synthetic_uniform
from sigp.instance import _cluster_points
from sigp.scenarios import GaussStream, gamma_draw, gen_durations

ROWS, COLS = 250, 160          # 40,000 draws per shape, one stream per row as in gen_durations
MEAN = 10.0


def _draws(shape, seed=7):
    return gen_durations(np.full(COLS, MEAN), ROWS, seed, mode="gamma", shape=shape).ravel()


def _erlang_cdf(x, k, theta):
    y = x / theta
    term, acc = 1.0, 1.0
    for i in range(1, k):
        term *= y / i
        acc += term
    return 1.0 - math.exp(-y) * acc


def _ks_stat(sample, cdf):
    xs = np.sort(sample)
    n = xs.size
    F = np.array([cdf(v) for v in xs])
    hi = np.arange(1, n + 1) / n - F
    lo = F - np.arange(0, n) / n
    return float(max(hi.max(), lo.max()))


def _ks_crit(n, alpha=0.001):
    return math.sqrt(-0.5 * math.log(alpha / 2.0)) / math.sqrt(n)


@pytest.mark.parametrize("shape", [0.5, 1.0, 2.0, 4.0])
def test_gamma_moments(shape):
    x = _draws(shape)
    theta = MEAN / shape
    n = x.size
    mu, var = shape * theta, shape * theta * theta
    assert abs(x.mean() - mu) < 5.0 * math.sqrt(var / n)
    # Var(s^2) ~ (mu4 - var^2)/n with mu4 = 3 k (k + 2) theta^4 for Gamma(k, theta)
    mu4 = 3.0 * shape * (shape + 2.0) * theta ** 4
    assert abs(x.var(ddof=1) - var) < 5.0 * math.sqrt((mu4 - var * var) / n)
    assert x.min() > 0.0 and np.all(np.isfinite(x))


@pytest.mark.parametrize("shape", [0.5, 1.0, 2.0, 4.0])
def test_gamma_ks_against_cdf(shape):
    x = _draws(shape, seed=11)
    theta = MEAN / shape
    D = stats.kstest(x, stats.gamma(shape, scale=theta).cdf).statistic
    assert D < _ks_crit(x.size), D


@pytest.mark.parametrize("k", [1, 2, 4])
def test_erlang_closed_form_and_exponential_sums(k):
    x = _draws(float(k), seed=13)
    theta = MEAN / k
    # the closed-form Erlang CDF agrees with scipy's Gamma CDF ...
    for v in (0.1, 1.0, 5.0, 10.0, 25.0, 60.0):
        assert _erlang_cdf(v, k, theta) == pytest.approx(stats.gamma(k, scale=theta).cdf(v), abs=1e-13)
    # ... and the sample is within the KS bound of it
    D = _ks_stat(x[:12000], lambda v: _erlang_cdf(v, k, theta))
    assert D < _ks_crit(12000), D
    # sums of k exponentials from an independent stream have the same law
    r = random.Random(2024 + k)
    sums = np.array([sum(-theta * math.log(1.0 - r.random()) for _ in range(k)) for _ in range(20000)])
    assert stats.ks_2samp(x, sums).pvalue > 0.001


def test_gamma_scale_equivariance():
    """Gamma(k, a*theta) = a * Gamma(k, theta) on the same stream: the
    sampler multiplies one standard draw by the scale."""
    for shape in (0.5, 2.0, 4.0):
        a = [gamma_draw(GaussStream(99), shape, 1.0) for _ in range(1)][0]
        b = gamma_draw(GaussStream(99), shape, 2.5)
        assert b == 2.5 * a or abs(b - 2.5 * a) <= 2 * np.spacing(b)


def test_gamma_rows_skip_zero_means():
    """Customers with expected duration <= 0 draw nothing and get 0, and
    the next customer's draw is the one it would have had without them."""
    e = np.array([10.0, 0.0, 10.0, -1.0, 10.0])
    rows = gen_durations(e, 3, 5, mode="gamma", shape=4.0)
    dense = gen_durations(np.array([10.0, 10.0, 10.0]), 3, 5, mode="gamma", shape=4.0)
    assert np.all(rows[:, [1, 3]] == 0.0)
    assert np.array_equal(rows[:, [0, 2, 4]], dense)


# ---------------------------------------------------------------------------
# synthetic generators

def _unclipped(inst):
    return (inst.open_t > 0.0) & (inst.close_t < inst.tmax)


def _coords(inst):
    xs = np.array([p.x for p in inst.pois])
    ys = np.array([p.y for p in inst.pois])
    return xs, ys


GENS = [
    ("uniform", lambda s: synthetic_uniform(n=100, m=2, seed=s), (10.0, 60.0)),
    ("clustered", lambda s: synthetic_clustered(n=100, m=3, seed=s), (30.0, 120.0)),
    ("mixed", lambda s: synthetic_mixed(n=288, m=4, seed=s), (20.0, 200.0)),
    ("uniform-wide", lambda s: synthetic_uniform(n=192, m=3, seed=s, width=(10.0, 200.0)), (10.0, 200.0)),
]


@pytest.mark.parametrize("name,gen,width", GENS, ids=[g[0] for g in GENS])
def test_generator_invariants(name, gen, width):
    for seed in (1, 2, 3):
        inst = gen(seed)
        xs, ys = _coords(inst)
        # integer coordinates in [0, 100]^2, depot first
        assert np.all(xs == np.round(xs)) and np.all(ys == np.round(ys))
        assert xs.min() >= 0 and ys.min() >= 0 and xs.max() <= 100 and ys.max() <= 100
        # depot window [0, tmax] = [0, 230]; expected service 10; scores U{1..50}
        assert inst.tmax == 230.0
        assert inst.pois[0].open_time == 0.0 and inst.pois[0].close_time == 230.0
        assert np.all(inst.durations == 10.0)
        sc = inst.scores
        assert np.all(sc == np.round(sc)) and sc.min() >= 1 and sc.max() <= 50
        # windows inside [0, tmax], no wider than the largest width; the
        # unclipped ones at least the smallest
        lo, hi = width
        w = inst.close_t - inst.open_t
        assert np.all(inst.open_t >= 0.0) and np.all(inst.close_t <= inst.tmax)
        assert np.all(w <= hi + 1e-9) and np.all(w > 0.0)
        un = _unclipped(inst)
        assert np.all(w[un] >= lo - 1e-9)
        # distances are sqrt(dx^2 + dy^2), symmetric bit for bit, zero diagonal
        D = np.asarray(inst.dist).reshape(inst.n + 1, inst.n + 1)
        assert np.array_equal(D, D.T) and np.all(np.diag(D) == 0.0)
        assert np.array_equal(inst.dist_to_depot, D[1:, 0])
        # window centres lie in [d(0,i), tmax - dur - d(i,0)] (collapsed to
        # d(0,i) when that is empty): visiting the customer straight from the
        # depot at t = 0 is feasible exactly when 2 d(0,i) + dur <= tmax
        d0 = inst.dist_to_depot
        arr = d0
        start = np.maximum(arr, inst.open_t)
        feasible = (arr <= inst.close_t) & (start + inst.durations + d0 <= inst.tmax)
        assert np.array_equal(feasible, 2.0 * d0 + inst.durations <= inst.tmax)
        c = 0.5 * (inst.open_t + inst.close_t)
        ok = 2.0 * d0 + inst.durations <= inst.tmax
        assert np.all(c[un & ok] >= d0[un & ok] - 1e-9)
        assert np.all(c[un & ok] <= (inst.tmax - inst.durations - d0)[un & ok] + 1e-9)


def test_generator_distributions():
    """Scores cover U{1..50} evenly; widths of unclipped windows are uniform
    on [lo, hi] (KS); uniform coordinates cover the square evenly."""
    scores = np.concatenate([synthetic_uniform(n=500, seed=s).scores for s in range(8)])
    counts = np.bincount(scores.astype(int), minlength=51)[1:]
    assert stats.chisquare(counts).pvalue > 0.001
    insts = [synthetic_uniform(n=500, seed=s, width=(10.0, 60.0)) for s in range(4)]
    w = np.concatenate([(i.close_t - i.open_t)[_unclipped(i)] for i in insts])
    assert stats.kstest(w, stats.uniform(10.0, 50.0).cdf).pvalue > 0.001
    xs = np.concatenate([_coords(i)[0][1:] for i in insts])
    assert stats.chisquare(np.bincount(xs.astype(int), minlength=101)).pvalue > 0.001


def test_cluster_spread():
    """One cluster: points are the rounded, clipped Gaussian around its
    centre, so away from the edges each axis has standard deviation
    sqrt(sigma^2 + 1/12) (rounding adds a uniform error of variance 1/12)."""
    for sigma in (3.0, 5.0, 8.0):
        for seed in range(40):
            pts = np.array(_cluster_points(random.Random(seed), 6000, 1, sigma), dtype=float)
            m = pts.mean(axis=0)
            if np.all(m > 4 * sigma) and np.all(m < 100 - 4 * sigma):
                break
        else:
            pytest.fail("no interior cluster centre in 40 seeds")
        want = math.sqrt(sigma * sigma + 1.0 / 12.0)
        se = want / math.sqrt(2 * pts.shape[0])
        for axis in (0, 1):
            assert abs(pts[:, axis].std(ddof=1) - want) < 5 * se
            # the counts per integer follow the rounded Gaussian's pmf
            # P(k) = Phi((k + 1/2 - mu) / sigma) - Phi((k - 1/2 - mu) / sigma)
            ks = np.arange(int(pts[:, axis].min()), int(pts[:, axis].max()) + 1)
            obs = np.array([(pts[:, axis] == k).sum() for k in ks], dtype=float)
            nd = stats.norm(m[axis], sigma)
            exp = pts.shape[0] * (nd.cdf(ks + 0.5) - nd.cdf(ks - 0.5))
            keep = exp >= 5
            obs_k, exp_k = obs[keep], exp[keep] * obs[keep].sum() / exp[keep].sum()
            assert stats.chisquare(obs_k, exp_k, ddof=1).pvalue > 0.001
        assert np.all(pts == np.round(pts)) and pts.min() >= 0 and pts.max() <= 100


def test_mixed_is_half_clustered():
    """synthetic_mixed: the first n // 2 customers are clustered (their
    nearest-neighbour distances are far shorter than the uniform half's)."""
    inst = synthetic_mixed(n=288, m=4, seed=1)
    xs, ys = _coords(inst)
    P = np.stack([xs[1:], ys[1:]], axis=1)

    def nn(block):
        d = np.sqrt(((block[:, None, :] - block[None, :, :]) ** 2).sum(-1))
        np.fill_diagonal(d, np.inf)
        return np.median(d.min(axis=1))

    assert nn(P[:144]) < 0.5 * nn(P[144:])
