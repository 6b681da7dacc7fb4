"""Distribution fills with P8_mt19937_64 adaptors as the engine, as the
reference's tests draw them (test/test.cpp:117 and on): the packet code takes
the adaptor as a scalar full-bit engine and packs its draws with RawbitsMaker
(arch/AVX/RandUtils.h:59-144, arch/SSE/RandUtils.h:22-67). Each GPU fill is
compared bit for bit with the oracle's, which is pinned to the libstdc++
mt19937_64 streams by tests/test_oracle_p8.py; fills of ragged sizes follow
one another so the operator() buffer position and the twists are carried
across calls, and the engine records must agree at the end."""
import numpy as np
import pytest

from test_gpu_fixed import assert_same

pytestmark = pytest.mark.gpu

SEEDS = [0, 5, 2 ** 40 + 17, 123456789] + [1000 + 16 * k for k in range(28)]

FLOAT_CASES = [
    ("uniform", -2.0, 3.0),
    ("std_normal", 0.0, 0.0),
    ("normal", 1.5, 0.25),
    ("exponential", 0.7, 0.0),
    ("cauchy", 0.5, 2.0),
    ("student_t", 4.5, 0.0),
    ("gamma", 0.4, 1.5),
    ("gamma", 3.7, 1.0),
    ("lognormal", 0.2, 0.7),
    ("weibull", 1.7, 0.8),
    ("extreme_value", 0.3, 1.2),
]
INT_CASES = [
    ("poisson", 2.0, 0.0),
    ("poisson", 30.0, 0.0),
    ("binomial", 10, 0.35),
    ("binomial", 200, 0.6),
]


def run_sequence(er, orc, dist, a, b, sizes):
    urng = er.P8_mt19937_64(seeds=SEEDS)
    ref = orc.p8_engines(SEEDS)
    for n in sizes:
        print(f"p8 {dist}({a}, {b}) count {n}", flush=True)
        got = er.p8_fill(dist, n, urng, a, b).cpu().numpy()
        assert_same(got, orc.p8_fill(ref, dist, n, a, b))
    np.testing.assert_array_equal(urng.snapshot(), ref.raw)


@pytest.mark.parametrize("dist,a,b", FLOAT_CASES, ids=[f"{c[0]}-{c[1]}" for c in FLOAT_CASES])
def test_p8_float_fills_match_oracle(er, orc, dist, a, b):
    # 8 floats = 3 scalars: every buffer position 0..7 occurs; 2000 and 5000
    # cross the engines' twists.
    run_sequence(er, orc, dist, a, b, [8, 24, 40, 8, 16, 2000, 8, 5000])


@pytest.mark.parametrize("dist,a,b", INT_CASES, ids=[f"{c[0]}-{c[1]}" for c in INT_CASES])
def test_p8_int_fills_match_oracle(er, orc, dist, a, b):
    run_sequence(er, orc, dist, a, b, [4, 12, 8, 4, 20, 1500, 4, 3000])


def test_p8_operator_stream_continues_across_fills(er, orc, mt_golden):
    """operator() from the buffer position: fills of 4 scalars leave half a
    refill; adaptor 0 (seed 0) is mt19937_64(0..7) interleaved."""
    urng = er.P8_mt19937_64(seeds=SEEDS)
    ref = orc.p8_engines(SEEDS)
    parts = []
    for n in (4, 12, 8, 20, 4, 4000, 4):
        got = er.p8_randbits_u64(n, urng).cpu().numpy()
        np.testing.assert_array_equal(got.ravel(), orc.p8_fill(ref, "u64", n))
        parts.append(got[0])
    s = np.concatenate(parts)
    streams = np.stack([mt_golden["streams"][j][: (s.size + 7) // 8] for j in range(8)], axis=1).ravel()
    np.testing.assert_array_equal(s, streams[: s.size])
    np.testing.assert_array_equal(urng.snapshot(), ref.raw)


def test_p8_mixed_distributions_share_the_buffer(er, orc):
    """Raw draws, float and int fills on the same adaptors in turn: each takes
    the scalars the previous one left in the operator() buffer."""
    urng = er.P8_mt19937_64(seeds=SEEDS)
    ref = orc.p8_engines(SEEDS)
    seq = [("u64", 4, 0, 0), ("std_normal", 8, 0, 0), ("poisson", 4, 3.0, 0), ("u64", 12, 0, 0),
           ("uniform", 16, 0.0, 1.0), ("binomial", 4, 20, 0.5), ("gamma", 24, 0.6, 2.0),
           ("u64", 8, 0, 0), ("poisson", 12, 40.0, 0), ("exponential", 8, 1.0, 0)]
    for dist, n, a, b in seq:
        print(f"p8 mixed {dist} count {n}", flush=True)
        if dist == "u64":
            got = er.p8_randbits_u64(n, urng).cpu().numpy()
            np.testing.assert_array_equal(got.ravel(), orc.p8_fill(ref, "u64", n))
        else:
            assert_same(er.p8_fill(dist, n, urng, a, b).cpu().numpy(), orc.p8_fill(ref, dist, n, a, b))
    np.testing.assert_array_equal(urng.snapshot(), ref.raw)


def test_p8_whole_refill_fills_reset_the_position(er, orc):
    """u32 fills start at a fresh refill: after 4 raw scalars the next 16
    words are refill 1, and the position is 0 again, so the next raw draw
    starts refill 2."""
    urng = er.P8_mt19937_64(seeds=SEEDS[:4])
    ref = orc.p8_engines(SEEDS[:4])
    all_raw = orc.p8_fill(ref, "u64", 24).reshape(4, 24)
    er.p8_randbits_u64(4, urng)
    w = er.p8_randbits_u32(16, urng).cpu().numpy()
    np.testing.assert_array_equal(w, all_raw[:, 8:16].copy().view(np.uint32))
    nxt = er.p8_randbits_u64(8, urng).cpu().numpy()
    np.testing.assert_array_equal(nxt, all_raw[:, 16:24])


def test_p8_fill_rejects(er):
    urng = er.P8_mt19937_64(seeds=[1, 2])
    with pytest.raises(ValueError):
        er.p8_fill("beta", 8, urng, 1.0, 2.0)
    with pytest.raises(er.InvalidArgument):
        er.p8_fill("normal", 12, urng, 0.0, 1.0)  # not a whole Packet8f
    with pytest.raises(er.InvalidArgument):
        er.p8_fill("poisson", 6, urng, 1.0)
