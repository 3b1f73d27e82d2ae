"""What the surrogate adds: trained on exactly simulated rules, the network's
choice of an eighth of a pool of fresh rules scores higher in exact
simulation than an eighth chosen at random, and its ranking of the pool
correlates with the exact one."""
import pytest

pytestmark = pytest.mark.gpu


def test_network_screen_beats_random_choice():
    from sigp.instance import synthetic_clustered
    from sigp.screening import Scorer, screening_gain, train_network
    sc = Scorer(synthetic_clustered(n=100, m=3, seed=1))
    net, norm, stats = train_network(sc, n_rules=6000, epochs=15, seed=1)
    print("network", stats, flush=True)
    assert stats["spearman"] > 0.7
    gain = screening_gain(sc, net, norm, pool=2048, seed=101)
    print("screening", gain, flush=True)
    assert gain["spearman_pool"] > 0.7
    assert gain["mean_kept_by_network"] > gain["mean_kept_at_random"] * 1.2
