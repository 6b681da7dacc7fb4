"""Gamma and Poisson vectorize-over-params fills at the full VoP configuration
size (2^24 elements, 4096 engines seeded 0,4,...,16380, parameters from
mt19937(2024): shape ~ logU(0.1, 50), mean ~ logU(0.01, 1e4)).

The whole GPU output passes the distribution checks of
test_oracle_vop_stats.py (PIT / randomised-PIT KS overall and per parameter
band, standardised moments), and sampled engine blocks are bit-exact against
the oracle.
"""
import numpy as np
import pytest

from test_oracle_vop_stats import check_gamma, check_poisson

pytestmark = pytest.mark.gpu

N_ENG = 4096
PER = 4096
N = N_ENG * PER  # 2^24
BLOCKS = (0, 1, 2047, 4095)


@pytest.fixture(scope="module")
def vop_params():
    rng = np.random.Generator(np.random.MT19937(2024))
    rng.uniform(-10, 10, N)  # normal means and sds come first in the configuration
    rng.uniform(0.1, 5, N)
    shape = np.exp(rng.uniform(np.log(0.1), np.log(50), N)).astype(np.float32)
    rng.integers(1, 1001, N)
    rng.uniform(0, 1, N)
    mean = np.exp(rng.uniform(np.log(0.01), np.log(1e4), N)).astype(np.float32)
    return shape, mean


def test_gamma_vop_full_size(er, orc, vop_params):
    import torch
    shape, _ = vop_params
    scale = np.ones(N, np.float32)
    urng = er.Vmt19937_64(0, n=N_ENG, stride=4)
    x = er.gamma_vop(urng, torch.from_numpy(shape).cuda(), torch.from_numpy(scale).cuda()).cpu().numpy()
    for g in BLOCKS:
        sl = slice(g * PER, (g + 1) * PER)
        ref = orc.Engines(seeds=[4 * g])
        exp = orc.fill_gamma_vop_f32(ref, shape[sl], scale[sl])
        np.testing.assert_array_equal(x[sl].view(np.uint32), exp.view(np.uint32))
    check_gamma(x, shape, scale)


def test_poisson_vop_full_size(er, orc, vop_params):
    import torch
    _, mean = vop_params
    urng = er.Vmt19937_64(0, n=N_ENG, stride=4)
    k = er.poisson_vop(urng, torch.from_numpy(mean).cuda()).cpu().numpy()
    for g in BLOCKS:
        sl = slice(g * PER, (g + 1) * PER)
        ref = orc.Engines(seeds=[4 * g])
        np.testing.assert_array_equal(k[sl], orc.fill_poisson_vop_i32(ref, mean[sl]))
    check_poisson(k, mean)
