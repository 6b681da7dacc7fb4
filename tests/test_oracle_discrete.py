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


# --- DiscreteGen<int32_t, double> (Discrete.h:606-779) ---

def _u64_stream(orc, seed, draws):
    """Tempered draws of Vmt19937_64(seed) as (draws, 4) lanes."""
    return orc.fill_u64(orc.Engines(seeds=[seed]), draws).reshape(draws, 4)


def _ur_f64(w):
    return ((w & np.uint64((1 << 52) - 1)) | np.uint64(1023 << 52)).view(np.float64) - 1.0


def _replay_discrete_d(orc, seed, w, n_packets):
    """The packetOp sequence restated from the raw stream: K < 16 takes one
    draw per packet; otherwise half() (new draw's low half, its high half on
    the next packet) and then one full draw per packet."""
    d = _u64_stream(orc, seed, 3 * n_packets + 2)
    w = np.asarray(w, np.float64)
    out, pos, cache = [], 0, None
    if w.size < 16:
        cdf = np.cumsum(w) / np.cumsum(w)[-1]
        for _ in range(n_packets):
            rx = _ur_f64(d[pos]); pos += 1
            out += [w.size - int((r < cdf).sum()) for r in rx]
        return np.array(out, np.int32)
    size = 1 << int(np.ceil(np.log2(w.size)))
    prob = np.zeros(size); alias = np.zeros(size, np.int32)
    orc.lib().orc_alias_table_f64(orc._p(w), orc.I64(w.size), orc._p(prob), orc._p(alias))
    for _ in range(n_packets):
        if cache is None:
            h, cache = d[pos][:2], d[pos][2:]
            pos += 1
        else:
            h, cache = cache, None
        words = np.array([int(h[0]) & 0xFFFFFFFF, int(h[0]) >> 32, int(h[1]) & 0xFFFFFFFF, int(h[1]) >> 32])
        idx = words & (size - 1)
        u = _ur_f64(d[pos]); pos += 1
        out += [int(i) if uu < prob[i] else int(alias[i]) for i, uu in zip(idx, u)]
    return np.array(out, np.int32)


@pytest.mark.parametrize("K", [3, 15, 16, 100, 1000])
def test_discrete_d_is_its_packet_sequence(orc, K):
    rng = np.random.default_rng(K + 5)
    w = rng.uniform(0, 1, K)
    for n_packets in (1, 2, 7, 64):
        got = orc.fill_discrete_d_i32(orc.Engines(seeds=[31]), 4 * n_packets, w)
        np.testing.assert_array_equal(got, _replay_discrete_d(orc, 31, w, n_packets))


def test_discrete_d_alias_table_is_the_double_form(orc):
    """The double table holds the same f values the float table rounds."""
    rng = np.random.default_rng(2)
    for K in (16, 100, 1000):
        w = rng.exponential(1.0, K)
        size = 1 << int(np.ceil(np.log2(K)))
        pd = np.zeros(size); ad = np.zeros(size, np.int32)
        pf = np.zeros(size, np.float32); af = np.zeros(size, np.int32)
        orc.lib().orc_alias_table_f64(orc._p(w), orc.I64(K), orc._p(pd), orc._p(ad))
        orc.lib().orc_alias_table_f32(orc._p(w), orc.I64(K), orc._p(pf), orc._p(af))
        np.testing.assert_array_equal(ad, af)
        np.testing.assert_array_equal(pd.astype(np.float32), pf)
        # what the table encodes: prob[s] of slot s, the rest to alias[s]
        p = np.zeros(size)
        for s in range(size):
            p[s] += pd[s]
            p[ad[s]] += 1 - pd[s]
        target = np.zeros(size); target[:K] = w / w.sum()
        assert np.abs(p / size - target).max() < 1e-12


def test_discrete_d_frequencies(orc):
    for w in ([1, 2, 3, 4], list(np.arange(1, 40, dtype=float))):
        w = np.asarray(w, dtype=np.float64)
        x = orc.fill_discrete_d_i32(orc.Engines(n=4, seed0=3), 1 << 16, w)
        counts = np.bincount(x, minlength=w.size)
        assert stats.chisquare(counts, w / w.sum() * x.size).pvalue > 1e-4
