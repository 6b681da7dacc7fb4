"""The oracle's samplers against the mathematics: queue order of the
rejection packets (hand-worked), and moments/KS against closed forms, in the
manner of the reference's accuracy checks (benchmark/accuracy.cpp:25-97,
test/test.cpp:472-494)."""
import math

import numpy as np
import pytest
from scipy import stats


def test_compress_queue_bypass_order(orc):
    # W=4: c0 accepts lanes {0,2}; c1 is fully accepted and is emitted at once,
    # ahead of the queued lanes; c2 accepts {1,2,3} and completes the queue.
    vals = np.arange(16, dtype=np.float32)
    masks = np.array([[1, 0, 1, 0], [1, 1, 1, 1], [0, 1, 1, 1], [1, 1, 0, 0]], dtype=np.uint8)
    out, used = orc.compress_stream(4, vals, masks, 2)
    np.testing.assert_array_equal(out[0], [4, 5, 6, 7])
    np.testing.assert_array_equal(out[1], [0, 2, 9, 10])
    assert used == 3  # c3 is never consumed


def test_compress_queue_overflow_carries(orc):
    # W=8: 5 + 5 accepted -> first packet = 5 from c0 + 3 from c1, rest = 2,
    # then a full candidate bypasses, then 6 more complete the second packet.
    vals = np.arange(32, dtype=np.float32)
    masks = np.zeros((4, 8), dtype=np.uint8)
    masks[0, :5] = 1
    masks[1, 3:] = 1
    masks[2, :] = 1
    masks[3, :6] = 1
    out, used = orc.compress_stream(8, vals, masks, 3)
    np.testing.assert_array_equal(out[0], [0, 1, 2, 3, 4, 11, 12, 13])
    np.testing.assert_array_equal(out[1], np.arange(16, 24))
    np.testing.assert_array_equal(out[2], [14, 15, 24, 25, 26, 27, 28, 29])
    assert used == 4


N_ENG = 8
PER = 16384


def sample(orc, fn, *args):
    e = orc.Engines(n=N_ENG, seed0=12345)
    return fn(e, PER, *args, threads=8)


def ks_ok(x, cdf):
    return stats.kstest(x, cdf).pvalue > 1e-4


@pytest.mark.parametrize("alpha,beta", [(0.2, 1.0), (0.5, 1.0), (1.0, 2.0), (3.0, 1.0), (3.7, 1.0),
                                        (10.5, 1.0), (25.0, 0.5)])
def test_gamma_all_regimes(orc, alpha, beta):
    x = sample(orc, orc.fill_gamma_f32, alpha, beta).astype(np.float64)
    assert np.all(x >= 0) and np.all(np.isfinite(x))
    assert ks_ok(x, stats.gamma(alpha, scale=beta).cdf)


def test_normal_exponential_cauchy_studentt(orc):
    assert ks_ok(sample(orc, orc.fill_normal_f32), stats.norm.cdf)
    assert ks_ok(sample(orc, orc.fill_normal_f32, 2.0, 3.0), stats.norm(2, 3).cdf)
    assert ks_ok(sample(orc, orc.fill_exponential_f32, 2.0), stats.expon(scale=0.5).cdf)
    assert ks_ok(sample(orc, orc.fill_cauchy_f32, 1.0, 2.0), stats.cauchy(1, 2).cdf)
    assert ks_ok(sample(orc, orc.fill_studentt_f32, 5.0), stats.t(5).cdf)
    assert ks_ok(sample(orc, orc.fill_uniform_f32), stats.uniform.cdf)


def _discrete_ok(x, dist, lo, hi):
    # chi-square over the bulk of the support
    ks = np.arange(lo, hi + 1)
    obs = np.array([(x == k).sum() for k in ks], dtype=np.float64)
    exp = dist.pmf(ks) * x.size
    keep = exp > 20
    chi = ((obs[keep] - exp[keep]) ** 2 / exp[keep]).sum()
    return stats.chi2(keep.sum() - 1).sf(chi) > 1e-5


@pytest.mark.parametrize("mean", [0.3, 1.0, 4.0, 11.5, 12.0, 16.0, 100.0, 3000.0])
def test_poisson(orc, mean):
    x = sample(orc, orc.fill_poisson_i32, mean)
    assert x.min() >= 0
    n = x.size
    assert abs(x.mean() - mean) < 5 * math.sqrt(mean / n) + 1e-9
    assert abs(x.var() / mean - 1) < 0.05
    sd = math.sqrt(mean)
    assert _discrete_ok(x, stats.poisson(mean), int(max(0, mean - 4 * sd)), int(mean + 4 * sd) + 1)


@pytest.mark.parametrize("n,p", [(20, 0.5), (10, 0.9), (50, 0.01), (100, 0.75), (1000, 0.3), (30, 0.02)])
def test_binomial(orc, n, p):
    x = sample(orc, orc.fill_binomial_i32, n, p)
    assert x.min() >= 0 and x.max() <= n or (n * min(p, 1 - p) < 1)
    m, v = n * p, n * p * (1 - p)
    assert abs(x.mean() - m) < 5 * math.sqrt(v / x.size) + 1e-9
    if n * min(p, 1 - p) >= 1:
        sd = math.sqrt(v)
        assert _discrete_ok(x, stats.binom(n, p), int(max(0, m - 4 * sd)), int(min(n, m + 4 * sd)))


def test_binomial_vop_means(orc):
    # binomialV row means/stdevs as in test/test.cpp:472-494
    rng = np.random.default_rng(3)
    n_per = 4 * 2048
    trials = np.repeat(rng.integers(25, 1000, N_ENG), n_per).astype(np.int32)
    p = np.repeat(rng.uniform(0.05, 0.95, N_ENG), n_per).astype(np.float32)
    e = orc.Engines(n=N_ENG, seed0=7)
    x = orc.fill_binomial_vop_i32(e, trials, p, threads=8).reshape(N_ENG, n_per)
    for g in range(N_ENG):
        nn, pp = trials[g * n_per], float(p[g * n_per])
        m, v = nn * pp, nn * pp * (1 - pp)
        assert abs(x[g].mean() - m) < 5 * math.sqrt(v / n_per)
        assert abs(x[g].var() / v - 1) < 0.15


@pytest.mark.parametrize("fill", ["fill_binomial_vop_i32", "fill_binomial_vopf_i32"])
def test_binomial_vop_randomised_pit(orc, fill):
    """Per-element (n_i, p_i) as in the VoP configuration (n ~ U{1..1000},
    p ~ U(0, 1)): the randomised PIT F(k-1) + V (F(k) - F(k-1)) is U(0, 1),
    for the int32 (Packet4i, half draws) and float (Packet8f, full draws)
    forms of BinomialVGen."""
    rng = np.random.default_rng(11)
    per = 4096
    n_eng = 32
    trials = rng.integers(1, 1001, n_eng * per).astype(np.int32)
    p = rng.uniform(0, 1, n_eng * per).astype(np.float32)
    e = orc.Engines(n=n_eng, seed0=99)
    k = getattr(orc, fill)(e, trials, p, threads=8).astype(np.float64)
    pp = p.astype(np.float64)
    lo = stats.binom.cdf(k - 1, trials, pp)
    hi = stats.binom.cdf(k, trials, pp)
    u = lo + np.random.default_rng(5).uniform(size=k.size) * (hi - lo)
    assert stats.kstest(u, "uniform").pvalue > 1e-4
    small = trials.astype(np.float64) * np.minimum(pp, 1 - pp) < 1  # the Knuth regime
    assert stats.kstest(u[small], "uniform").pvalue > 1e-4
    assert stats.kstest(u[~small], "uniform").pvalue > 1e-4
