"""Vectorize-over-params gamma/Poisson and batched requests (segment programs)
against the oracle and against the equivalent sequence of single fills."""
import numpy as np
import pytest

from test_gpu_fixed import assert_same, flat

pytestmark = pytest.mark.gpu

SEEDS = [17, 2 ** 35 + 3, 16380, 5]


def test_gamma_vop_matches_oracle(er, orc):
    import torch
    rng = np.random.default_rng(7)
    per = 2048
    n = len(SEEDS) * per
    alpha = np.exp(rng.uniform(np.log(0.1), np.log(50), n)).astype(np.float32)  # logU(0.1, 50)
    alpha[::97] = 1.0
    beta = rng.uniform(0.5, 2, n).astype(np.float32)
    urng = er.Vmt19937_64(seeds=SEEDS)
    ref = orc.Engines(seeds=SEEDS)
    for _ in range(2):
        out = er.gamma_vop(urng, torch.from_numpy(alpha).cuda(), torch.from_numpy(beta).cuda())
        assert_same(out.cpu().numpy(), orc.fill_gamma_vop_f32(ref, alpha, beta))
        np.testing.assert_array_equal(urng.snapshot(), ref.raw)


def test_poisson_vop_matches_oracle(er, orc):
    import torch
    rng = np.random.default_rng(8)
    per = 2048
    n = len(SEEDS) * per
    mean = np.exp(rng.uniform(np.log(0.01), np.log(1e4), n)).astype(np.float32)  # logU(0.01, 1e4)
    urng = er.Vmt19937_64(seeds=SEEDS)
    ref = orc.Engines(seeds=SEEDS)
    for _ in range(2):
        out = er.poisson_vop(urng, torch.from_numpy(mean).cuda())
        assert_same(out.cpu().numpy(), orc.fill_poisson_vop_i32(ref, mean))
        np.testing.assert_array_equal(urng.snapshot(), ref.raw)


def test_poisson_vop_long_knuth_chains(er, orc):
    """Means in [5, 12): every lane runs the Knuth product chain, often over
    more than one 16-draw batch and across twists; edge means 0.01, 11.99, 12."""
    import torch
    rng = np.random.default_rng(81)
    per = 4096
    n = len(SEEDS) * per
    mean = rng.uniform(5.0, 12.0, n).astype(np.float32)
    mean[::53] = np.float32(11.99)
    mean[::71] = np.float32(12.0)
    mean[::89] = np.float32(0.01)
    urng = er.Vmt19937_64(seeds=SEEDS)
    ref = orc.Engines(seeds=SEEDS)
    out = er.poisson_vop(urng, torch.from_numpy(mean).cuda())
    assert_same(out.cpu().numpy(), orc.fill_poisson_vop_i32(ref, mean))
    np.testing.assert_array_equal(urng.snapshot(), ref.raw)


def test_requests_equal_sequential_fills(er, orc):
    """A mixed batch on 3 engines equals the same fills issued one by one."""
    import torch
    dev = "cuda"
    specs = [  # (engine, dist, count, a, b)
        (0, "normal", 24, 1.0, 2.0), (1, "gamma", 16, 0.5, 1.0), (0, "poisson", 12, 20.0, 0),
        (2, "binomial", 8, 30, 0.4), (1, "uniform", 40, -1.0, 1.0), (0, "gamma", 8, 3.7, 1.0),
        (2, "std_normal", 8, 0, 0), (1, "poisson", 20, 3.0, 0), (2, "binomial", 4, 10, 0.7),
        (0, "exponential", 8, 2.0, 0), (1, "cauchy", 8, 0.0, 1.0), (2, "student_t", 16, 5.0, 0),
        (0, "gamma", 16, 4.0, 0.5), (2, "discard", 7, 0, 0), (2, "u64", 8, 0, 0),
        (1, "binomial", 8, 1000, 0.001), (0, "uniform_f64", 8, 0.0, 1.0),
    ]
    dtypes = {"poisson": torch.int32, "binomial": torch.int32, "u64": torch.uint64,
              "uniform_f64": torch.float64}
    reqs, outs = [], []
    for eng, dist, count, a, b in specs:
        o = None if dist == "discard" else torch.empty(count, dtype=dtypes.get(dist, torch.float32), device=dev)
        outs.append(o)
        reqs.append({"dist": dist, "engine": eng, "count": count, "out": o, "a": a, "b": b})
    seeds = [100, 200, 300]
    urng = er.Vmt19937_64(seeds=seeds)
    er.run_requests(urng, reqs)
    ref = orc.Engines(seeds=seeds)
    for (eng, dist, count, a, b), o in zip(specs, outs):
        e1 = ref.subset(eng, 1)
        if dist == "normal":
            exp = orc.fill_normal_f32(e1, count, a, b)
        elif dist == "std_normal":
            exp = orc.fill_normal_f32(e1, count)
        elif dist == "gamma":
            exp = orc.fill_gamma_f32(e1, count, a, b)
        elif dist == "poisson":
            exp = orc.fill_poisson_i32(e1, count, a)
        elif dist == "binomial":
            exp = orc.fill_binomial_i32(e1, count, int(a), b)
        elif dist == "uniform":
            exp = orc.fill_uniform_f32(e1, count, a, b)
        elif dist == "uniform_f64":
            exp = orc.fill_uniform_f64(e1, count, a, b)
        elif dist == "exponential":
            exp = orc.fill_exponential_f32(e1, count, a)
        elif dist == "cauchy":
            exp = orc.fill_cauchy_f32(e1, count, a, b)
        elif dist == "student_t":
            exp = orc.fill_studentt_f32(e1, count, a)
        elif dist == "u64":
            exp = orc.fill_u64(e1, count // 4)
        elif dist == "discard":
            orc.fill_u64(e1, count)
            exp = None
        ref.raw[eng * orc.STATE_BYTES:(eng + 1) * orc.STATE_BYTES] = e1.raw
        if exp is not None:
            assert_same(o.cpu().numpy(), exp)
    np.testing.assert_array_equal(urng.snapshot(), ref.raw)
    assert urng.status() == 0


def test_requests_reject_bad_arguments(er):
    import torch
    urng = er.Vmt19937_64(1, n=2)
    o = torch.empty(8, device="cuda")
    with pytest.raises(er.InvalidArgument):
        er.run_requests(urng, [{"dist": "normal", "engine": 5, "count": 8, "out": o}])
    with pytest.raises(er.InvalidArgument):
        er.run_requests(urng, [{"dist": "normal", "engine": 0, "count": 7, "out": o}])
    with pytest.raises(er.InvalidArgument):
        er.run_requests(urng, [{"dist": "gamma", "engine": 0, "count": 8, "out": o, "a": -2.0}])
