"""The mixed scalar/array overloads normal(urng, a[], b), normal(urng, a, b[])
(NormalExp.h:1304-1335) and binomial(urng, trials[], p), binomial(urng,
trials, p[]) (Discrete.h:1852-1897): the reference passes the scalar as a
constant array of the same shape, so each must equal the all-array overload,
and the oracle, with that array, bit for bit."""
import numpy as np
import pytest

from test_gpu_fixed import assert_same

pytestmark = pytest.mark.gpu

SEEDS = [9, 2 ** 33 + 1, 77]
PER = 4096


def test_binomial_mixed_overloads(er, orc):
    import torch
    rng = np.random.default_rng(3)
    n = len(SEEDS) * PER
    trials = rng.integers(1, 1001, n).astype(np.int32)
    p = rng.uniform(0, 1, n).astype(np.float32)
    for tr, pr, t_arg, p_arg in (
            (trials, np.full(n, 0.3, np.float32), torch.from_numpy(trials).cuda(), 0.3),
            (np.full(n, 400, np.int32), p, 400, torch.from_numpy(p).cuda()),
            (np.full(n, 30, np.int32), p * np.float32(0.02), 30, torch.from_numpy(p * np.float32(0.02)).cuda())):
        a = er.Vmt19937_64(seeds=SEEDS)
        b = er.Vmt19937_64(seeds=SEEDS)
        ref = orc.Engines(seeds=SEEDS)
        got = er.binomial_vop(a, t_arg, p_arg).cpu().numpy()
        full = er.binomial_vop(b, torch.from_numpy(tr).cuda(), torch.from_numpy(pr).cuda()).cpu().numpy()
        np.testing.assert_array_equal(got, full)
        assert_same(got, orc.fill_binomial_vop_i32(ref, tr, pr))
        np.testing.assert_array_equal(a.snapshot(), ref.raw)


def test_normal_mixed_overloads(er, orc):
    import torch
    rng = np.random.default_rng(4)
    n = len(SEEDS) * PER
    mean = rng.uniform(-10, 10, n).astype(np.float32)
    sd = rng.uniform(0.1, 5, n).astype(np.float32)
    for mu, s, m_arg, s_arg in (
            (mean, np.full(n, 2.5, np.float32), torch.from_numpy(mean).cuda(), 2.5),
            (np.full(n, -1.0, np.float32), sd, -1.0, torch.from_numpy(sd).cuda())):
        a = er.Vmt19937_64(seeds=SEEDS)
        ref = orc.Engines(seeds=SEEDS)
        got = er.normal_vop(a, m_arg, s_arg).cpu().numpy()
        assert_same(got, orc.fill_normal_vop_f32(ref, mu, s))
        np.testing.assert_array_equal(a.snapshot(), ref.raw)


def test_binomial_float_form_matches_oracle(er, orc):
    """BinomialVGen<float> (Discrete.h:1196-1331): Packet8f lanes on full
    draws, per-element (trials, p) including the edges (n = 0, p = 0 and 1,
    the Knuth regime), then the scalar-p form on the continuing engines."""
    import torch
    rng = np.random.default_rng(8)
    n = len(SEEDS) * PER
    trials = rng.integers(1, 1001, n).astype(np.int32)
    p = rng.uniform(0, 1, n).astype(np.float32)
    trials[::5] = rng.integers(1, 30, trials[::5].size)
    p[::7] = np.float32(0.003)
    trials[::11] = 0
    p[::13] = np.float32(1.0)
    p[::17] = np.float32(0.0)
    a = er.Vmt19937_64(seeds=SEEDS)
    ref = orc.Engines(seeds=SEEDS)
    t_dev = torch.from_numpy(trials).cuda()
    got = er.binomial_vop_float(a, t_dev, torch.from_numpy(p).cuda()).cpu().numpy()
    assert_same(got, orc.fill_binomial_vopf_i32(ref, trials, p))
    got = er.binomial_vop_float(a, t_dev, 0.37).cpu().numpy()
    assert_same(got, orc.fill_binomial_vopf_i32(ref, trials, np.full(n, 0.37, np.float32)))
    np.testing.assert_array_equal(a.snapshot(), ref.raw)
    assert a.status() == 0
