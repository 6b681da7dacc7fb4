"""Distribution checks that pin the oracle's gamma and Poisson
vectorize-over-params samplers to the mathematics.

These two VoP samplers have no counterpart in the reference, so the exact
stream replays (test_oracle_stream_pins.py) only show that the oracle does
what its own lockstep rules say. Here the outputs are checked against the
target distributions with per-element parameters drawn as in the VoP
configuration (shape ~ logU(0.1, 50), mean ~ logU(0.01, 1e4)):

- gamma: the probability integral transform u_i = F(x_i; alpha_i, beta_i) is
  U(0, 1) (KS, overall and per shape band, so a regime-specific mistake
  cannot hide in the mixture), and the standardised residuals
  (x_i - alpha_i beta_i) / (sqrt(alpha_i) beta_i) have mean 0 and variance 1;
- Poisson: the randomised PIT u_i = F(k_i - 1) + V_i (F(k_i) - F(k_i - 1)),
  V_i ~ U(0, 1), is U(0, 1) (KS, overall and per mean band), chi-square of
  the counts for constant means on both sides of the Knuth/tangent switch at
  12, and the residuals (k_i - mu_i) / sqrt(mu_i) have mean 0 and variance 1.

A dropped term, a wrong sign in an accept test or a swapped parameter moves
these statistics far outside their sampling bands (the planted-mistake tests
at the end show it for three such mistakes).
"""
import numpy as np
import pytest
from scipy import stats

N_ENG = 256
PER = 4096
N = N_ENG * PER
P_MIN = 1e-4


def vop_params(seed=2024):
    rng = np.random.default_rng(seed)
    alpha = np.exp(rng.uniform(np.log(0.1), np.log(50), N)).astype(np.float32)
    beta = rng.uniform(0.5, 2.0, N).astype(np.float32)
    mean = np.exp(rng.uniform(np.log(0.01), np.log(1e4), N)).astype(np.float32)
    return alpha, beta, mean


def gamma_pit(x, alpha, beta):
    return stats.gamma.cdf(np.asarray(x, np.float64), np.asarray(alpha, np.float64),
                           scale=np.asarray(beta, np.float64))


def poisson_rpit(k, mean, seed=7):
    k = np.asarray(k, np.float64)
    mu = np.asarray(mean, np.float64)
    v = np.random.default_rng(seed).uniform(size=k.size)
    lo = stats.poisson.cdf(k - 1, mu)
    hi = stats.poisson.cdf(k, mu)
    return lo + v * (hi - lo)


def ks_p(u):
    return stats.kstest(u, "uniform").pvalue


def check_gamma(x, alpha, beta, bands=(0.1, 0.5, 1.0, 1.0 + 1e-6, 5.0, 50.0 + 1e-3)):
    x = np.asarray(x, np.float64)
    assert np.all(np.isfinite(x)) and np.all(x >= 0)
    u = gamma_pit(x, alpha, beta)
    assert ks_p(u) > P_MIN
    a = np.asarray(alpha, np.float64)
    for lo, hi in zip(bands[:-1], bands[1:]):
        sel = (a >= lo) & (a < hi)
        if sel.sum() > 500:
            assert ks_p(u[sel]) > P_MIN, (lo, hi)
    b = np.asarray(beta, np.float64)
    z = (x - a * b) / (np.sqrt(a) * b)
    n = z.size
    assert abs(z.mean()) < 5 / np.sqrt(n)
    # the 4th moment of the standardised gamma is 3 + 6/alpha, at most 63 here
    assert abs(z.var() - 1) < 5 * np.sqrt(np.mean(2 + 6 / a) / n)


def check_poisson(k, mean, bands=(0.01, 1.0, 12.0, 100.0, 1e4 + 1)):
    k = np.asarray(k)
    assert np.all(k >= 0)
    u = poisson_rpit(k, mean)
    assert ks_p(u) > P_MIN
    mu = np.asarray(mean, np.float64)
    for lo, hi in zip(bands[:-1], bands[1:]):
        sel = (mu >= lo) & (mu < hi)
        if sel.sum() > 500:
            assert ks_p(u[sel]) > P_MIN, (lo, hi)
    z = (k - mu) / np.sqrt(mu)
    n = z.size
    assert abs(z.mean()) < 5 / np.sqrt(n)
    # the 4th moment of the standardised Poisson is 3 + 1/mu
    assert abs(z.var() - 1) < 5 * np.sqrt(np.mean(2 + 1 / mu) / n)


def test_gamma_vop_distribution(orc):
    alpha, beta, _ = vop_params()
    e = orc.Engines(n=N_ENG, seed0=99)
    x = orc.fill_gamma_vop_f32(e, alpha, beta, threads=8)
    check_gamma(x, alpha, beta)


def test_poisson_vop_distribution(orc):
    _, _, mean = vop_params()
    e = orc.Engines(n=N_ENG, seed0=77)
    k = orc.fill_poisson_vop_i32(e, mean, threads=8)
    check_poisson(k, mean)


@pytest.mark.parametrize("a", [0.1, 0.5, 0.999, 1.0, 1.001, 3.7, 50.0])
def test_gamma_vop_constant_shape(orc, a):
    n_eng, per = 16, 8192
    alpha = np.full(n_eng * per, a, np.float32)
    beta = np.full(n_eng * per, 1.5, np.float32)
    e = orc.Engines(n=n_eng, seed0=5)
    x = orc.fill_gamma_vop_f32(e, alpha, beta, threads=8).astype(np.float64)
    assert stats.kstest(x, stats.gamma(float(np.float32(a)), scale=1.5).cdf).pvalue > P_MIN


def _chi2_p(k, mu):
    lo = int(max(0, np.floor(mu - 6 * np.sqrt(mu) - 2)))
    hi = int(np.ceil(mu + 6 * np.sqrt(mu) + 3))
    ks = np.arange(lo, hi + 1)
    obs = np.array([(k == j).sum() for j in ks], dtype=np.float64)
    exp = stats.poisson.pmf(ks, mu) * k.size
    keep = exp > 20
    chi = ((obs[keep] - exp[keep]) ** 2 / exp[keep]).sum()
    return stats.chi2(max(1, keep.sum() - 1)).sf(chi)


@pytest.mark.parametrize("mu", [0.01, 0.5, 5.0, 11.9, 12.0, 12.1, 100.0, 1e4])
def test_poisson_vop_constant_mean(orc, mu):
    n_eng, per = 16, 8192
    mean = np.full(n_eng * per, mu, np.float32)
    e = orc.Engines(n=n_eng, seed0=11)
    k = orc.fill_poisson_vop_i32(e, mean, threads=8)
    assert _chi2_p(k, float(np.float32(mu))) > 1e-5


# --- the checks reject plausible mistakes (planted in the oracle's outputs) ---

def test_checks_reject_a_swapped_gamma_scale(orc):
    alpha, beta, _ = vop_params()
    e = orc.Engines(n=N_ENG, seed0=99)
    x = orc.fill_gamma_vop_f32(e, alpha, np.ones_like(beta), threads=8)
    with pytest.raises(AssertionError):
        check_gamma(x * alpha / alpha.mean(), alpha, beta)  # scale not applied per element
    with pytest.raises(AssertionError):
        check_gamma(x, alpha, beta)  # beta dropped


def test_checks_reject_a_shape_off_by_one_in_one_regime(orc):
    alpha, beta, _ = vop_params()
    e = orc.Engines(n=N_ENG, seed0=99)
    shifted = np.where(alpha < 1, alpha + 0.1, alpha).astype(np.float32)
    x = orc.fill_gamma_vop_f32(e, shifted, beta, threads=8)
    with pytest.raises(AssertionError):
        check_gamma(x, alpha, beta)


def test_checks_reject_a_poisson_mean_rounded(orc):
    _, _, mean = vop_params()
    e = orc.Engines(n=N_ENG, seed0=77)
    k = orc.fill_poisson_vop_i32(e, np.floor(mean) + 1, threads=8)
    with pytest.raises(AssertionError):
        check_poisson(k, mean)
