"""GPU parity of the rejection (CompressMask) and lockstep samplers: gamma in
its GS and tangent regimes, Poisson, binomial and BinomialVGen. Integer
outputs and engine records must be identical to the oracle's; sequences mix
half-draw consumers with full-draw fills so that fills start with a cached
half draw and at odd block offsets."""
import numpy as np
import pytest

from test_gpu_fixed import assert_same, flat

pytestmark = pytest.mark.gpu

SEEDS = [5, 2 ** 33 + 1, 16380, 42, 999]


@pytest.mark.parametrize("alpha,beta", [(0.2, 1.0), (0.5, 1.0), (0.99, 2.0), (3.7, 1.0), (10.5, 1.0),
                                        (25.0, 0.5), (1.5, 3.0)])
def test_gamma_reject_matches_oracle(er, orc, alpha, beta):
    urng = er.Vmt19937_64(seeds=SEEDS)
    ref = orc.Engines(seeds=SEEDS)
    for n in (8, 16, 2504, 4096, 8):
        out = er.gamma(n, 1, urng, alpha, beta)
        assert_same(flat(out), orc.fill_gamma_f32(ref, n, alpha, beta))
    np.testing.assert_array_equal(urng.snapshot(), ref.raw)
    assert urng.status() == 0


@pytest.mark.parametrize("alpha,beta", [(0.5, 1.0), (3.7, 2.0)])
def test_gamma_reject_one_wave_per_engine_matches_oracle(er, orc, alpha, beta):
    """Above 1024 engines the rejection fill runs one wave per engine (up to
    1024, a four-wave workgroup per engine: test_gamma_reject_matches_oracle);
    sampled engine blocks and records against the oracle."""
    n_eng, per = 1028, 2504
    urng = er.Vmt19937_64(77, n=n_eng, stride=4)
    out = er.gamma(per, 1, urng, alpha, beta).reshape(n_eng, per).cpu().numpy()
    snap = urng.snapshot().reshape(n_eng, -1)
    for g in (0, 1, 513, n_eng - 1):
        ref = orc.Engines(seeds=[77 + 4 * g])
        assert_same(out[g], orc.fill_gamma_f32(ref, per, alpha, beta))
        np.testing.assert_array_equal(snap[g], ref.raw)
    assert urng.status() == 0


@pytest.mark.parametrize("mean", [0.3, 1.0, 4.0, 11.5, 12.0, 16.0, 100.0, 3000.0])
def test_poisson_matches_oracle(er, orc, mean):
    urng = er.Vmt19937_64(seeds=SEEDS)
    ref = orc.Engines(seeds=SEEDS)
    for n in (4, 12, 2500, 4):
        out = er.poisson(n, 1, urng, mean)
        assert_same(flat(out), orc.fill_poisson_i32(ref, n, mean))
        np.testing.assert_array_equal(urng.snapshot(), ref.raw)
        # a full-draw fill in between leaves any cached half draw in place
        u = er.uniform_real(8, 1, urng)
        assert_same(flat(u), orc.fill_uniform_f32(ref, 8))
    np.testing.assert_array_equal(urng.snapshot(), ref.raw)


@pytest.mark.parametrize("trials,p", [(20, 0.5), (10, 0.9), (1, 0.3), (0, 0.5), (24, 1.0),
                                      (50, 0.01), (30, 0.98), (100, 0.75), (1000, 0.3), (25, 0.5)])
def test_binomial_matches_oracle(er, orc, trials, p):
    urng = er.Vmt19937_64(seeds=SEEDS)
    ref = orc.Engines(seeds=SEEDS)
    for n in (4, 8, 1252, 4):
        out = er.binomial(n, 1, urng, trials, p)
        assert_same(flat(out), orc.fill_binomial_i32(ref, n, trials, p))
        np.testing.assert_array_equal(urng.snapshot(), ref.raw)
        z = er.normal(8, 1, urng)
        assert_same(flat(z), orc.fill_normal_f32(ref, 8))


def test_binomial_vop_matches_oracle(er, orc):
    """Config-3 parameter ranges: n ~ U{1..1000}, p ~ U(0,1), int32 bit-exact."""
    import torch
    rng = np.random.default_rng(2024)
    per = 4096
    n = len(SEEDS) * per
    trials = rng.integers(1, 1001, n).astype(np.int32)
    p = rng.uniform(0, 1, n).astype(np.float32)
    urng = er.Vmt19937_64(seeds=SEEDS)
    ref = orc.Engines(seeds=SEEDS)
    for _ in range(2):
        out = er.binomial_vop(urng, torch.from_numpy(trials).cuda(), torch.from_numpy(p).cuda())
        assert_same(out.cpu().numpy(), orc.fill_binomial_vop_i32(ref, trials, p))
        np.testing.assert_array_equal(urng.snapshot(), ref.raw)


def test_binomial_vop_edge_parameters(er, orc):
    import torch
    vals_n = np.array([0, 1, 1, 24, 25, 1000, 1000, 7], dtype=np.int32)
    vals_p = np.array([0.5, 0.0, 1.0, 0.5, 0.04, 1e-4, 0.9999, 0.5], dtype=np.float32)
    trials = np.tile(vals_n, 64)
    p = np.tile(vals_p, 64)
    urng = er.Vmt19937_64(seeds=[3])
    ref = orc.Engines(seeds=[3])
    out = er.binomial_vop(urng, torch.from_numpy(trials).cuda(), torch.from_numpy(p).cuda())
    assert_same(out.cpu().numpy(), orc.fill_binomial_vop_i32(ref, trials, p))


@pytest.mark.parametrize("mean", [0.01, 2.5, 7.0, 11.9])
def test_poisson_knuth_long_fills(er, orc, mean):
    """Long Knuth-regime fills: many twists, windows that chain several
    packets and packets that run past a window."""
    urng = er.Vmt19937_64(seeds=SEEDS)
    ref = orc.Engines(seeds=SEEDS)
    for n in (16384, 36):
        out = er.poisson(n, 1, urng, mean)
        assert_same(flat(out), orc.fill_poisson_i32(ref, n, mean, threads=8))
    np.testing.assert_array_equal(urng.snapshot(), ref.raw)
    # BinomialGen's Poisson regime (n >= 25, n*min(p,1-p) < 1) with the flip
    out = er.binomial(4096, 1, urng, 400, 0.998)
    assert_same(flat(out), orc.fill_binomial_i32(ref, 4096, 400, 0.998, threads=8))
    np.testing.assert_array_equal(urng.snapshot(), ref.raw)


@pytest.mark.parametrize("mean", [2.2, 3.9, 4.1, 5.0, 9.99])
def test_poisson_knuth_prefix_many_windows(er, orc, mean):
    """Means on both sides of the prefix-path switch (mean 4), over
    enough packets that lanes within the fixed-point margin (recomputed with
    the exact loop) and windows crossing twists occur many times."""
    urng = er.Vmt19937_64(seeds=SEEDS)
    ref = orc.Engines(seeds=SEEDS)
    for n in (1 << 18, 4, 8):
        out = er.poisson(n, 1, urng, mean)
        assert_same(flat(out), orc.fill_poisson_i32(ref, n, mean, threads=8))
        np.testing.assert_array_equal(urng.snapshot(), ref.raw)


@pytest.mark.parametrize("kind,a,b", [("poisson", 1.03e6, 0), ("poisson", 1048000.0, 0),
                                      ("poisson", 3.0e6, 0), ("binomial", 2 ** 21 + 5, 0.5),
                                      ("binomial", 3000000, 0.4), ("binomial", 2100000, 0.97)])
def test_tangent_beyond_lgamma_table(er, orc, kind, a, b):
    """Tangent candidates whose plgamma arguments straddle or exceed the 2^20
    entries of the device table (computed exactly there)."""
    import torch
    urng = er.Vmt19937_64(seeds=SEEDS)
    ref = orc.Engines(seeds=SEEDS)
    n = 2048
    if kind == "poisson":
        assert_same(flat(er.poisson(n, 1, urng, a)), orc.fill_poisson_i32(ref, n, a))
    else:
        assert_same(flat(er.binomial(n, 1, urng, a, b)), orc.fill_binomial_i32(ref, n, a, b))
    np.testing.assert_array_equal(urng.snapshot(), ref.raw)
    # the same parameters per lane in the VoP forms
    m = len(SEEDS) * 512
    if kind == "poisson":
        mean = np.full(m, a, np.float32)
        mean[::3] = np.float32(20.0)
        out = er.poisson_vop(urng, torch.from_numpy(mean).cuda())
        assert_same(out.cpu().numpy(), orc.fill_poisson_vop_i32(ref, mean))
    else:
        trials = np.full(m, a, np.int32)
        trials[::3] = 100
        pr = np.full(m, b, np.float32)
        out = er.binomial_vop(urng, torch.from_numpy(trials).cuda(), torch.from_numpy(pr).cuda())
        assert_same(out.cpu().numpy(), orc.fill_binomial_vop_i32(ref, trials, pr))
    np.testing.assert_array_equal(urng.snapshot(), ref.raw)


@pytest.mark.parametrize("trials,p", [(1, 0.5), (3, 0.7), (10, 0.5), (25, 0.3), (50, 0.05), (19, 0.9)])
def test_negative_binomial_matches_oracle(er, orc, trials, p):
    """NegativeBinomialGen: gamma means on half draws (exponential, integer
    and tangent regimes), then Knuth or tangent Poisson sharing the gamma's
    queue; odd packet counts leave cached halves between fills."""
    urng = er.Vmt19937_64(seeds=SEEDS)
    ref = orc.Engines(seeds=SEEDS)
    for n in (4, 12, 1004, 4):
        assert_same(flat(er.negative_binomial(n, 1, urng, trials, p)),
                    orc.fill_negative_binomial_i32(ref, n, trials, p))
        np.testing.assert_array_equal(urng.snapshot(), ref.raw)
        assert_same(flat(er.uniform_real(8, 1, urng)), orc.fill_uniform_f32(ref, 8))
    assert urng.status() == 0
