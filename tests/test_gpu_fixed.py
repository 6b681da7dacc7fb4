"""GPU parity of the engine and the fixed-consumption fills against the oracle.

Integer/bit outputs must be identical; float fills are compared bit for bit
too (the kernels follow the oracle's op sequence exactly).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(x):
    x = np.asarray(x)
    return x.view(np.uint32) if x.dtype == np.float32 else x.view(np.uint64)


def assert_same(got, exp):
    got = np.asarray(got).ravel()
    exp = np.asarray(exp).ravel()
    assert got.shape == exp.shape
    bad = np.nonzero(bits(got) != bits(exp))[0]
    assert bad.size == 0, f"{bad.size} mismatches, first at {bad[:5]}: {got[bad[:5]]} vs {exp[bad[:5]]}"


def flat(t):
    """Column-major matrix view (or flat buffer) back to the linear fill order."""
    if t.dim() == 2:
        return t.t().contiguous().view(-1).cpu().numpy()
    if t.dim() == 3:
        return t.transpose(1, 2).contiguous().view(-1).cpu().numpy()
    return t.contiguous().view(-1).cpu().numpy()


def test_raw_stream_matches_std(er, mt_golden):
    urng = er.Vmt19937_64(42)
    d = er.randbits_u64(1000, urng).cpu().numpy().reshape(1000, 4)
    for lane in range(4):
        np.testing.assert_array_equal(d[:, lane], mt_golden["streams"][42 + lane])


def test_raw_stream_known_answer(er, mt_golden):
    urng = er.Vmt19937_64()  # default seed 5489
    d = er.randbits_u64(10000, urng).cpu().numpy().reshape(10000, 4)
    assert int(d[9999, 0]) == 9981545732273789042


def test_engine_records_match_oracle(er, orc):
    urng = er.Vmt19937_64(7, n=3, stride=5)
    ref = orc.Engines(seeds=[7, 12, 17])
    np.testing.assert_array_equal(urng.snapshot(), ref.raw)
    er.randbits_u64(700, urng)
    orc.fill_u64(ref, 700)
    np.testing.assert_array_equal(urng.snapshot(), ref.raw)


def test_config1_sequence(er, orc):
    """Vmt19937_64(42): 2^20 raw u64, then uniformReal and normal 1024x1024."""
    urng = er.Vmt19937_64(42)
    ref = orc.Engines(seeds=[42])
    raw = er.randbits_u64(1 << 18, urng).cpu().numpy()
    assert_same(raw, orc.fill_u64(ref, 1 << 18))
    u = er.uniform_real(1024, 1024, urng)
    assert_same(flat(u), orc.fill_uniform_f32(ref, 1 << 20))
    z = er.normal(1024, 1024, urng)
    assert_same(flat(z), orc.fill_normal_f32(ref, 1 << 20))
    np.testing.assert_array_equal(urng.snapshot(), ref.raw)


CASES = [
    ("uniform", lambda er, n, g: er.uniform_real(n, 1, g), lambda o, e, n: o.fill_uniform_f32(e, n)),
    ("uniform_ab", lambda er, n, g: er.uniform_real(n, 1, g, -3.0, 5.5),
     lambda o, e, n: o.fill_uniform_f32(e, n, -3.0, 5.5)),
    ("uniform_f64", lambda er, n, g: er.uniform_real_f64(n, 1, g), lambda o, e, n: o.fill_uniform_f64(e, n)),
    ("std_normal", lambda er, n, g: er.normal(n, 1, g), lambda o, e, n: o.fill_normal_f32(e, n)),
    ("std_normal_f64", lambda er, n, g: er.normal_f64(n, 1, g), lambda o, e, n: o.fill_normal_f64(e, n)),
    ("normal_f64", lambda er, n, g: er.normal_f64(n, 1, g, -1.5, 0.3),
     lambda o, e, n: o.fill_normal_f64(e, n, -1.5, 0.3)),
    ("exponential_f64", lambda er, n, g: er.exponential_f64(n, 1, g, 0.7),
     lambda o, e, n: o.fill_exponential_f64(e, n, 0.7)),
    ("cauchy_f64", lambda er, n, g: er.cauchy_f64(n, 1, g, 0.5, 2.0),
     lambda o, e, n: o.fill_cauchy_f64(e, n, 0.5, 2.0)),
    ("normal", lambda er, n, g: er.normal(n, 1, g, 2.0, 3.0), lambda o, e, n: o.fill_normal_f32(e, n, 2.0, 3.0)),
    ("exponential", lambda er, n, g: er.exponential(n, 1, g, 1.5),
     lambda o, e, n: o.fill_exponential_f32(e, n, 1.5)),
    ("cauchy", lambda er, n, g: er.cauchy(n, 1, g, 0.5, 2.0), lambda o, e, n: o.fill_cauchy_f32(e, n, 0.5, 2.0)),
    ("student_t", lambda er, n, g: er.student_t(n, 1, g, 5.0), lambda o, e, n: o.fill_studentt_f32(e, n, 5.0)),
    ("gamma_1", lambda er, n, g: er.gamma(n, 1, g, 1.0, 2.0), lambda o, e, n: o.fill_gamma_f32(e, n, 1.0, 2.0)),
    ("gamma_3", lambda er, n, g: er.gamma(n, 1, g, 3.0, 1.5), lambda o, e, n: o.fill_gamma_f32(e, n, 3.0, 1.5)),
    ("gamma_19", lambda er, n, g: er.gamma(n, 1, g, 19.0, 1.0), lambda o, e, n: o.fill_gamma_f32(e, n, 19.0, 1.0)),
    ("chi_squared_4", lambda er, n, g: er.chi_squared(n, 1, g, 4.0), lambda o, e, n: o.fill_gamma_f32(e, n, 2.0, 2.0)),
    ("lognormal", lambda er, n, g: er.lognormal(n, 1, g, 0.5, 0.75),
     lambda o, e, n: o.fill_lognormal_f32(e, n, 0.5, 0.75)),
    ("weibull", lambda er, n, g: er.weibull(n, 1, g, 1.7, 2.5), lambda o, e, n: o.fill_weibull_f32(e, n, 1.7, 2.5)),
    ("extreme_value", lambda er, n, g: er.extreme_value(n, 1, g, 1.0, 3.0),
     lambda o, e, n: o.fill_extreme_value_f32(e, n, 1.0, 3.0)),
    ("bernoulli", lambda er, n, g: er.bernoulli(n, 1, g, 0.3), lambda o, e, n: o.fill_bernoulli_f32(e, n, 0.3)),
    ("bernoulli_1", lambda er, n, g: er.bernoulli(n, 1, g, 1.0), lambda o, e, n: o.fill_bernoulli_f32(e, n, 1.0)),
    ("balanced", lambda er, n, g: er.balanced(n, 1, g), lambda o, e, n: o.fill_balanced_f32(e, n)),
    ("balanced_ab", lambda er, n, g: er.balanced(n, 1, g, -3.0, 5.5),
     lambda o, e, n: o.fill_balanced_f32(e, n, -3.0, 5.5)),
]


@pytest.mark.parametrize("name,gpu,cpu", CASES, ids=[c[0] for c in CASES])
def test_fill_matches_oracle(er, orc, name, gpu, cpu):
    # 5 engines; sizes chosen so that engines start fills at odd block offsets
    # and normal fills end on an unused cos half-packet (8 mod 16).
    seeds = [3, 1000, 2 ** 40 + 5, 16380, 77]
    urng = er.Vmt19937_64(seeds=seeds)
    ref = orc.Engines(seeds=seeds)
    for n in (8, 24, 312 * 8 + 8, 4000, 8):
        if name.endswith("_f64"):
            n = n // 2  # 4 doubles per draw; 4 mod 8 ends normal fills on a sine packet
        out = gpu(er, n, urng)
        exp = cpu(orc, ref, n)
        assert_same(flat(out), exp)
    np.testing.assert_array_equal(urng.snapshot(), ref.raw)


def test_normal_vop_matches_oracle(er, orc):
    import torch
    seeds = [11, 12, 13]
    per = 3000 * 8
    rng = np.random.default_rng(5)
    mean = rng.uniform(-10, 10, 3 * per).astype(np.float32)
    sd = rng.uniform(0.1, 5, 3 * per).astype(np.float32)
    urng = er.Vmt19937_64(seeds=seeds)
    ref = orc.Engines(seeds=seeds)
    er.randbits_u64(3, urng)
    orc.fill_u64(ref, 3)
    out = er.normal_vop(urng, torch.from_numpy(mean).cuda(), torch.from_numpy(sd).cuda())
    assert_same(out.cpu().numpy(), orc.fill_normal_vop_f32(ref, mean, sd))


def test_snapshot_restore_copy_discard(er, orc):
    urng = er.Vmt19937_64(99, n=4)
    er.randbits_u64(100, urng)
    snap = urng.snapshot()
    a = er.normal(512, 1, urng)
    urng.restore(snap)
    b = er.normal(512, 1, urng)
    assert_same(flat(a), flat(b))
    c = urng.copy()
    x = er.uniform_real(256, 1, urng)
    y = er.uniform_real(256, 1, c)
    assert_same(flat(x), flat(y))
    # discard(n) == drawing n packets
    d = urng.copy()
    urng.discard(1000)
    er.randbits_u64(1000, d)
    np.testing.assert_array_equal(urng.snapshot(), d.snapshot())


def test_views_fill_their_engines(er, orc):
    urng = er.Vmt19937_64(0, n=6)
    ref = orc.Engines(n=6, seed0=0)
    v = urng.view(2, 3)
    out = er.uniform_real(64, 1, v)
    exp = orc.fill_uniform_f32(ref.subset(2, 3), 64)
    assert_same(flat(out), exp)


def test_invalid_arguments_raise(er):
    urng = er.Vmt19937_64(1)
    with pytest.raises(er.InvalidArgument):
        er.normal(7, 1, urng)  # not a multiple of the 8-float packet
    with pytest.raises(er.InvalidArgument):
        er.gamma(8, 1, urng, -1.0, 1.0)


def test_full_size_normal_sampled_parity_and_moments(er, orc):
    """2^28 normals from 4096 engines seeded 0,4,...,16380 (16384 MT lanes
    seeded 0..16383): four engine blocks bit-exact vs the oracle, and the
    whole output's mean/variance/kurtosis within sampling error."""
    import torch
    n_eng, per = 4096, 65536
    urng = er.Vmt19937_64(0, n=n_eng, stride=4)
    out = er.normal(per, 1, urng)  # (n, per, 1)
    out = out.reshape(n_eng, per)
    for g in (0, 1, 2047, 4095):
        ref = orc.Engines(seeds=[4 * g])
        assert_same(out[g].cpu().numpy(), orc.fill_normal_f32(ref, per))
    x = out.reshape(-1).double()
    m = x.mean().item()
    v = x.var().item()
    k = ((x - m) ** 4).mean().item() / v ** 2
    n = x.numel()
    assert abs(m) < 5 / n ** 0.5
    assert abs(v - 1) < 5 * (2 / n) ** 0.5
    assert abs(k - 3) < 5 * (24 / n) ** 0.5
    assert torch.isfinite(out).all()


def test_p8_adaptor_streams(er, orc, mt_golden):
    """P8_mt19937_64: out[8k + j] = mt19937_64(seed + j)[k]; the u32 and
    uniform_real() fills continue the same refill sequence."""
    urng = er.P8_mt19937_64(seeds=[0, 1000])
    assert urng.adaptors == 2 and len(urng) == 4
    d = er.p8_randbits_u64(8 * 1000, urng).cpu().numpy()
    a0 = d[0].reshape(1000, 8)
    for j in range(8):
        np.testing.assert_array_equal(a0[:, j], mt_golden["streams"][j])
    ref = orc.Engines(seeds=[1000, 1004])
    np.testing.assert_array_equal(d[1], orc.fill_p8_u64(ref, 1000))
    u32 = er.p8_randbits_u32(16 * 64, urng).cpu().numpy()
    uf = er.p8_uniform_real(16 * 64, urng).cpu().numpy()
    w = orc.fill_p8_u64(ref, 128).view(np.uint32)  # 16 words per refill
    np.testing.assert_array_equal(u32[1], w[:16 * 64])
    np.testing.assert_array_equal(uf[1], ((w[16 * 64:] & 0x7FFFFF) / 2.0 ** 23).astype(np.float32))
    snap = urng.snapshot().reshape(4, -1)
    np.testing.assert_array_equal(snap[2:].ravel(), ref.raw)


def test_p8_rejects_bad_sizes(er):
    urng = er.P8_mt19937_64(5)
    with pytest.raises(er.InvalidArgument):
        er.p8_randbits_u64(10, urng)
    with pytest.raises(er.InvalidArgument):
        er.p8_uniform_real(8, urng)
    odd = er.Vmt19937_64(1, n=3)
    with pytest.raises(er.InvalidArgument):
        er.p8_randbits_u64(8, odd)


GEN_CASES = [
    ("std_normal", 3, 0.0, 0.0, [8, 24, 40, 8, 16]),
    ("lognormal", 16, 0.2, 0.7, [8, 8, 24]),
    ("normal", 4, 2.0, 0.5, [24, 8, 16, 8]),
    ("gamma", 9, 0.4, 1.5, [8, 16, 40, 8, 1000]),
    ("gamma", 9, 3.7, 1.0, [16, 8, 24, 2000]),
    ("poisson", 10, 30.0, 0.0, [4, 12, 20, 4, 2500]),
    ("poisson", 10, 2.5, 0.0, [4, 8, 12]),
    ("binomial", 11, 100, 0.3, [4, 12, 8, 1500]),
]


@pytest.mark.parametrize("name,dist,a,b,sizes", GEN_CASES)
def test_generator_objects_match_oracle(er, orc, name, dist, a, b, sizes):
    seeds = [3, 77, 2 ** 40 + 5, 16380]
    urng = er.Vmt19937_64(seeds=seeds)
    ref = orc.Engines(seeds=seeds)
    gen = er.Generator(name, a, b, len(seeds))
    rgen = orc.Generators(dist, a, b, len(seeds))
    for n in sizes:
        got = gen.generate(n, urng).cpu().numpy()
        assert_same(got, rgen.fill(ref, n))
    np.testing.assert_array_equal(urng.snapshot(), ref.raw)
    assert urng.status() == 0


def test_offsets_past_2_31_elements(er, orc):
    """Two engines of 2^30 + 8 floats: engine 1's block ends past element
    2^31, so its tail checks 64-bit output indexing."""
    import torch
    per = (1 << 30) + 8
    urng = er.Vmt19937_64(seeds=[3, 11])
    out = torch.empty(2 * per, dtype=torch.float32, device="cuda")
    er.uniform_real(per, 1, urng, out=out)
    tail = out[-64:].cpu().numpy()
    ref = orc.Engines(seeds=[11])
    orc.discard(ref, per // 8 - 8, threads=1)
    assert_same(tail, orc.fill_uniform_f32(ref, 64))
    np.testing.assert_array_equal(urng.snapshot()[10016:], ref.raw)
    del out


def test_many_engines_and_empty_fills(er, orc):
    """2^20 engines in one launch; zero-size fills leave engines untouched."""
    n = 1 << 20
    urng = er.Vmt19937_64(7, n=n, stride=4)
    before = urng.view(n - 2, 2).snapshot()
    er.normal(0, 1, urng)
    er.poisson(0, 1, urng, 30.0)
    np.testing.assert_array_equal(urng.view(n - 2, 2).snapshot(), before)
    x = er.normal(16, 1, urng).reshape(n, 16)
    ref = orc.Engines(seeds=[7 + 4 * (n - 1)])
    assert_same(x[n - 1].cpu().numpy(), orc.fill_normal_f32(ref, 16))


HALF_CASES = [
    ("bernoulli_i32", lambda er, n, g: er.bernoulli(n, 1, g, 0.3, dtype=__import__("torch").int32),
     lambda o, e, n: o.fill_bernoulli_i32(e, n, 0.3)),
    ("geometric", lambda er, n, g: er.geometric(n, 1, g, 0.3), lambda o, e, n: o.fill_geometric_i32(e, n, 0.3)),
    ("geometric_tiny_p", lambda er, n, g: er.geometric(n, 1, g, 1e-9),
     lambda o, e, n: o.fill_geometric_i32(e, n, 1e-9)),
    ("randbits_i32", lambda er, n, g: er.randbits_i32(n, 1, g), lambda o, e, n: o.fill_randbits_i32(e, n)),
]


@pytest.mark.parametrize("name,gpu,cpu", HALF_CASES, ids=[c[0] for c in HALF_CASES])
def test_half_draw_fill_matches_oracle(er, orc, name, gpu, cpu):
    """Packet4i fills through half(): odd packet counts leave the upper half
    cached for the next half-draw fill, and a full-draw fill in between leaves
    it in place; engine records must match after every step."""
    seeds = [3, 1000, 2 ** 40 + 5, 16380, 77]
    urng = er.Vmt19937_64(seeds=seeds)
    ref = orc.Engines(seeds=seeds)
    for n in (4, 12, 312 * 8 + 4, 2500, 4):
        assert_same(flat(gpu(er, n, urng)), cpu(orc, ref, n))
        np.testing.assert_array_equal(urng.snapshot(), ref.raw)
        assert_same(flat(er.uniform_real(8, 1, urng)), orc.fill_uniform_f32(ref, 8))


VOP_CASES = [
    ("uniform", "uniform_real_vop", 2), ("exponential", "exponential_vop", 1), ("cauchy", "cauchy_vop", 2),
    ("weibull", "weibull_vop", 2), ("extreme_value", "extreme_value_vop", 2), ("studentt", "student_t_vop", 1),
    ("lognormal", "lognormal_vop", 2),
]


@pytest.mark.parametrize("kind,fn,nparams", VOP_CASES, ids=[c[0] for c in VOP_CASES])
def test_vop_fill_matches_oracle(er, orc, kind, fn, nparams):
    import torch
    seeds = [11, 12, 13]
    per = 3000 * 8 + 8  # lognormal: ends on an unused cos half packet
    rng = np.random.default_rng(len(kind))
    a = rng.uniform(0.3, 4, 3 * per).astype(np.float32)
    b = rng.uniform(0.3, 4, 3 * per).astype(np.float32)
    urng = er.Vmt19937_64(seeds=seeds)
    ref = orc.Engines(seeds=seeds)
    er.randbits_u64(3, urng)
    orc.fill_u64(ref, 3)
    args = [torch.from_numpy(a).cuda()] + ([torch.from_numpy(b).cuda()] if nparams == 2 else [])
    out = getattr(er, fn)(urng, *args)
    assert_same(out.cpu().numpy(), orc.fill_vop_f32(ref, kind, a, b if nparams == 2 else None))
    np.testing.assert_array_equal(urng.snapshot(), ref.raw)


def test_balanced_and_bernoulli_vop_match_oracle(er, orc):
    import torch
    seeds = [21, 22]
    per = 4096
    rng = np.random.default_rng(3)
    a = rng.uniform(-5, 5, 2 * per).astype(np.float32)
    b = (a + rng.uniform(0.1, 5, 2 * per)).astype(np.float32)
    p = rng.uniform(0, 1, 2 * per).astype(np.float32)
    p[:8] = [0, 1, 0.5, 1e-9, 1 - 1e-7, 0.25, 1, 0]
    urng = er.Vmt19937_64(seeds=seeds)
    ref = orc.Engines(seeds=seeds)
    got = er.balanced_vop(urng, torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
    assert_same(got.cpu().numpy(), orc.fill_balanced_vop_f32(ref, a, b))
    got = er.bernoulli_vop(urng, torch.from_numpy(p).cuda())
    assert_same(got.cpu().numpy(), orc.fill_bernoulli_vop_f32(ref, p))
    np.testing.assert_array_equal(urng.snapshot(), ref.raw)


@pytest.mark.parametrize("lo,hi", [(0, 9), (-3, 4), (0, 2), (10, 1000), (0, 7), (7, 7), (-100, 2 ** 30),
                                   (0, 2 ** 31 - 1), (0, 2 ** 20 + 1), (-2 ** 31, -2 ** 31 + 5)])
def test_uniform_int_matches_oracle(er, orc, lo, hi):
    """UniformIntGen: constant, power-of-two and CompressMask ranges; odd
    packet counts leave cached halves, and a full-draw fill in between."""
    seeds = [3, 1000, 2 ** 40 + 5, 16380, 77]
    urng = er.Vmt19937_64(seeds=seeds)
    ref = orc.Engines(seeds=seeds)
    for n in (4, 12, 312 * 8 + 4, 20000, 4):
        assert_same(flat(er.uniform_int(n, 1, urng, lo, hi)), orc.fill_uniform_int_i32(ref, n, lo, hi))
        np.testing.assert_array_equal(urng.snapshot(), ref.raw)
        assert_same(flat(er.uniform_real(8, 1, urng)), orc.fill_uniform_f32(ref, 8))


def test_uniform_int_rejects_bad_ranges(er):
    urng = er.Vmt19937_64(1)
    with pytest.raises(er.ErdError):
        er.uniform_int(8, 1, urng, 5, 4)
    with pytest.raises(er.ErdError):
        er.uniform_int(8, 1, urng, -2 ** 31, 2 ** 31 - 1)


@pytest.mark.parametrize("a,b,fisher", [(0.4, 0.7, False), (2.0, 3.0, False), (0.5, 4.5, False), (1.0, 1.0, False),
                                        (3.3, 0.2, False), (7.0, 2.5, True), (0.6, 1.2, True),
                                        # integer shapes: fixed draws per packet (a + b up to 38)
                                        (1.0, 4.0, False), (19.0, 19.0, False), (5.0, 1.0, False),
                                        (4.0, 6.0, True), (2.0, 38.0, True)])
def test_beta_fisher_f_match_oracle(er, orc, a, b, fisher):
    """BetaGen in both regimes (Johnk; GammaGen pairs in GS, integer,
    exponential and tangent regimes) and FisherFGen."""
    seeds = [3, 1000, 2 ** 40 + 5]
    urng = er.Vmt19937_64(seeds=seeds)
    ref = orc.Engines(seeds=seeds)
    for n in (8, 16, 2504, 8):
        if fisher:
            got, exp = er.fisher_f(n, 1, urng, a, b), orc.fill_fisher_f_f32(ref, n, a, b)
        else:
            got, exp = er.beta(n, 1, urng, a, b), orc.fill_beta_f32(ref, n, a, b)
        assert_same(flat(got), exp)
        np.testing.assert_array_equal(urng.snapshot(), ref.raw)
    assert urng.status() == 0
