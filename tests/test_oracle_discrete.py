"""Oracle pins for DiscreteGen / AliasMethod / MultinomialGen.

The alias table is checked against what it must encode: for every category,
(acceptance mass of its own slot + mass redirected to it by alias slots) /
table size = its probability, up to the int32 quantisation of prob (2^-31 per
slot). Category frequencies and multinomial moments follow test/test.cpp:
773-925 and test_mv.cpp:181-201."""
import numpy as np
import pytest
from scipy import stats


def implied_probs(prob, alias, K):
    size = prob.size
    p = np.zeros(size)
    for s in range(size):
        acc = prob[s] / 2.0 ** 31
        p[s] += acc
        p[alias[s]] += 1 - acc
    return p / size


@pytest.mark.parametrize("K", [16, 17, 33, 100, 1000, 4096])
def test_alias_table_encodes_weights(orc, K):
    rng = np.random.default_rng(K)
    w = rng.exponential(1.0, K)
    w[rng.integers(0, K, K // 10)] = 0.0  # some empty categories
    prob, alias = orc.alias_table(w)
    p = implied_probs(prob, alias, K)
    target = np.zeros(prob.size)
    target[:K] = w / w.sum()
    assert np.abs(p - target).max() < 2.0 ** -30 * 4
    assert np.all(alias < K) and np.all(prob >= 0)


def test_discrete_frequencies(orc):
    for w in ([1, 2, 3, 4], list(np.arange(1, 40, dtype=float))):
        w = np.asarray(w, dtype=np.float64)
        e = orc.Engines(n=4, seed0=3)
        x = orc.fill_discrete_i32(e, 1 << 16, w)
        counts = np.bincount(x, minlength=w.size)
        chi = stats.chisquare(counts, w / w.sum() * x.size)
        assert chi.pvalue > 1e-4


@pytest.mark.parametrize("trials,K", [(20, 4), (99, 10), (1000, 4), (5000, 100), (50, 1000)])
def test_multinomial_moments(orc, trials, K):
    rng = np.random.default_rng(trials + K)
    w = rng.uniform(0.1, 1, K)
    p = w / w.sum()
    e = orc.Engines(n=1, seed0=9)
    x = orc.multinomial(e, trials, [w], 2048)[0].astype(np.float64)
    assert np.all(x.sum(axis=0) == trials)
    sd = np.sqrt(trials * p * (1 - p))
    assert np.all(np.abs(x.mean(axis=1) - trials * p) < 5 * sd / np.sqrt(2048) + 1e-9)
