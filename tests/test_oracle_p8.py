"""The oracle's P8_mt19937_64 scalar-engine feed, pinned to the libstdc++
mt19937_64 golden streams: the adaptor's scalar stream is
S[8r + j] = mt19937_64(seed + j)[r]; packet distributions take it through
RawbitsMaker<..., scalar_fullbit> (arch/AVX/RandUtils.h:59-144,
arch/SSE/RandUtils.h:22-67) in Clang's argument order. Checked here:
operator() itself (with the buffer position carried across fills), the
Packet8f uniform_real packing (rawbits_34) by its closed form, Box-Muller on
those uniforms, the Packet4i packing through Poisson's Knuth loop, and the
distributions' shapes."""
import numpy as np
import pytest
from scipy import stats

SEED = 0  # adaptor lanes mt19937_64(0..7): golden streams 0..7


def scalar_stream(mt_golden, n):
    s = np.stack([mt_golden["streams"][j][: (n + 7) // 8] for j in range(8)], axis=1)
    return s.reshape(-1)[:n]


def test_p8_operator_stream_with_buffer_position(orc, mt_golden):
    e = orc.p8_engines([SEED])
    parts = [orc.p8_fill(e, "u64", n) for n in (4, 12, 8, 20, 4)]  # positions 4, 0, 0, 4, 0
    np.testing.assert_array_equal(np.concatenate(parts), scalar_stream(mt_golden, 48))
    assert int(e.raw.view(np.uint64).reshape(2, -1)[0, -1]) == 0


def rawbits_34_floats(s3):
    """Packet8f uniform_real of scalars (a, b, c): words [a.lo, a.hi, b.lo,
    X1, b.hi, c.lo, c.hi, X2], float = (word & 0x7FFFFF) / 2^23."""
    a, b, c = (int(x) for x in s3)
    w = [a & 0xFFFFFFFF, a >> 32, b & 0xFFFFFFFF, 0, b >> 32, c & 0xFFFFFFFF, c >> 32, 0]
    top = lambda x: x >> 24
    w[3] = top(w[0]) | top(w[1]) << 8 | top(w[2]) << 16 | top(w[2]) << 24
    w[7] = top(w[4]) | top(w[5]) << 8 | top(w[6]) << 16 | top(w[6]) << 24
    return np.array([(x & 0x7FFFFF) / 2.0 ** 23 for x in w], np.float32)


def test_p8_uniform_real_packing(orc, mt_golden):
    e = orc.p8_engines([SEED])
    got = orc.p8_fill(e, "uniform", 8 * 40, 0.0, 1.0).reshape(40, 8)
    s = scalar_stream(mt_golden, 3 * 40)
    for v in range(40):
        np.testing.assert_array_equal(got[v], rawbits_34_floats(s[3 * v: 3 * v + 3]))
    # 120 scalars = 15 whole refills: the buffer is empty again
    assert int(e.raw.view(np.uint64).reshape(2, -1)[0, -1]) == 0


def test_p8_normal_is_box_muller_of_the_packed_uniforms(orc, mt_golden):
    e = orc.p8_engines([SEED])
    z = orc.p8_fill(e, "std_normal", 16 * 30).reshape(30, 2, 8).astype(np.float64)
    s = scalar_stream(mt_golden, 6 * 30)
    for m in range(30):
        u1 = 1.0 - rawbits_34_floats(s[6 * m: 6 * m + 3]).astype(np.float64)
        u2 = rawbits_34_floats(s[6 * m + 3: 6 * m + 6]).astype(np.float64)
        r = np.sqrt(-2 * np.log(u1))
        np.testing.assert_allclose(z[m, 0], r * np.sin(2 * np.pi * u2), atol=4e-6, rtol=4e-6)
        np.testing.assert_allclose(z[m, 1], r * np.cos(2 * np.pi * u2), atol=4e-6, rtol=4e-6)
    # 180 scalars: 22 refills and 4 scalars of the 23rd
    assert int(e.raw.view(np.uint64).reshape(2, -1)[0, -1]) == 4


def test_p8_packet4_packing_through_knuth(orc, mt_golden):
    """Poisson(2) (Knuth, Packet4f uniform_real via the Packet4i packing): a
    packet multiplies its lanes' uniforms until every lane is <= e^-2; the
    half draw of scalars (first, second) puts the second in u64 lane 0."""
    e = orc.p8_engines([SEED])
    got = orc.p8_fill(e, "poisson", 4 * 50, 2.0)
    s = scalar_stream(mt_golden, 4000)
    thr = np.float32(np.exp(-2.0))
    pos, out = 0, []
    for _ in range(50):
        val = np.ones(4, np.float32)
        res = np.zeros(4, np.int32)
        while True:
            first, second = int(s[pos]), int(s[pos + 1])
            pos += 2
            lanes = [second & 0xFFFFFFFF, second >> 32, first & 0xFFFFFFFF, first >> 32]
            u = np.array([(w & 0x7FFFFF) / 2.0 ** 23 for w in lanes], np.float32)
            val = (val * u).astype(np.float32)
            c = thr < val
            if not c.any():
                break
            res += c
        out += list(res)
    np.testing.assert_array_equal(got, np.array(out, np.int32))


@pytest.mark.parametrize("dist,a,b,cdf", [
    ("uniform", 0.0, 1.0, stats.uniform.cdf),
    ("normal", 2.0, 3.0, stats.norm(2, 3).cdf),
    ("exponential", 2.0, 0.0, stats.expon(scale=0.5).cdf),
    ("gamma", 0.5, 1.0, stats.gamma(0.5).cdf),
    ("gamma", 3.7, 2.0, stats.gamma(3.7, scale=2).cdf),
    ("cauchy", 0.0, 1.0, stats.cauchy.cdf),
    ("student_t", 5.0, 0.0, stats.t(5).cdf),
])
def test_p8_float_distributions(orc, dist, a, b, cdf):
    e = orc.p8_engines([11, 19, 27, 35])
    x = orc.p8_fill(e, dist, 1 << 15, a, b).astype(np.float64)
    assert stats.kstest(x, cdf).pvalue > 1e-4


@pytest.mark.parametrize("dist,a,b", [("poisson", 0.7, 0), ("poisson", 9.0, 0), ("poisson", 40.0, 0),
                                      ("binomial", 10, 0.3), ("binomial", 200, 0.4), ("binomial", 500, 0.001)])
def test_p8_int_distributions(orc, dist, a, b):
    e = orc.p8_engines([11, 19])
    k = orc.p8_fill(e, dist, 1 << 15, a, b)
    d = stats.poisson(a) if dist == "poisson" else stats.binom(int(a), b)
    vals = np.arange(max(0, int(d.mean() - 8 * d.std() - 2)), int(d.mean() + 8 * d.std() + 3))
    obs = np.array([(k == v).sum() for v in vals], np.float64)
    exp = d.pmf(vals) * k.size
    keep = exp > 20
    chi = ((obs[keep] - exp[keep]) ** 2 / exp[keep]).sum()
    assert stats.chi2(keep.sum() - 1).sf(chi) > 1e-5
