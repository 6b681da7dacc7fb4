"""Pins of the oracle's samplers against the raw engine stream.

Each test draws the engine's raw u64 stream (itself pinned to std::mt19937_64)
and recomputes the samples independently of the oracle's C++ code. The
recomputation is written from the reference's packetOp text in Python: which
draws feed which output slot, the accept tests, and the CompressMask queue.

- Fixed-consumption samplers (normal, exponential, Cauchy, Student-t) are
  checked in float64 within a tolerance.
- Integer and rejection samplers are checked exactly, with float32 numpy
  arithmetic for the control flow. For the transcendental bits they use the
  oracle's elementwise plog/pexp/psincos/plgamma, which test_oracle_math pins
  to libm.

A dropped term, a swapped draw or a wrong queue order fails here.
"""
import math

import numpy as np
import pytest

f32 = np.float32
EPS8 = f32(2.0 ** -26)  # nonzero_uniform_real: U + FLT_EPSILON / 8


def full_words(orc, e, draws):
    """(n_engines, draws, 8) u32 words: float j of a draw = word j."""
    raw = orc.fill_u64(e.copy(), draws)
    return raw.view(np.uint32).reshape(len(e), draws, 8)


def half_words(orc, e, halves):
    """(n_engines, halves, 4) u32 words of the half() sequence."""
    out = []
    for g in range(len(e)):
        raw = orc.fill_half_u64(e.subset(g, 1).copy(), halves)
        out.append(raw.view(np.uint32).reshape(halves, 4))
    return np.stack(out)


def ur(w):
    return ((w & 0x7FFFFF).astype(np.float64) / 2.0 ** 23).astype(f32)


def test_normal_is_box_muller_of_the_stream(orc):
    e = orc.Engines(n=3, seed0=77)
    per = 1024
    got = orc.fill_normal_f32(e.copy(), per).reshape(3, per // 16, 2, 8)
    w = full_words(orc, e, per // 8).reshape(3, per // 16, 2, 8)
    u1 = 1.0 - ur(w[:, :, 0, :]).astype(np.float64)   # call 2m -> 1 - U
    u2 = ur(w[:, :, 1, :]).astype(np.float64)         # call 2m+1 -> theta
    r = np.sqrt(-2 * np.log(u1))
    want_sin = r * np.sin(2 * math.pi * u2)
    want_cos = r * np.cos(2 * math.pi * u2)
    tol = 4e-6 * (1 + np.abs(want_sin))
    assert np.all(np.abs(got[:, :, 0, :] - want_sin) <= tol)
    assert np.all(np.abs(got[:, :, 1, :] - want_cos) <= 4e-6 * (1 + np.abs(want_cos)))


def test_exponential_cauchy_studentt_of_the_stream(orc):
    e = orc.Engines(n=2, seed0=5)
    per = 512
    w = full_words(orc, e, 2 * per // 8)
    u = ur(w[:, :per // 8, :]).astype(np.float64).reshape(2, per)
    ex = orc.fill_exponential_f32(e.copy(), per, 2.0).reshape(2, per)
    np.testing.assert_allclose(ex, -np.log(1 - u) / 2.0, rtol=3e-6, atol=1e-7)
    ca = orc.fill_cauchy_f32(e.copy(), per, 1.0, 3.0).reshape(2, per)
    # tan is ill-conditioned near the poles: take the float32 argument
    x = (f32(math.pi) * (ur(w[:, :per // 8, :]).reshape(2, per) - f32(0.5))).astype(np.float64)
    want = 1.0 + 3.0 * np.tan(x)
    np.testing.assert_allclose(ca, want, rtol=2e-5, atol=2e-5)
    # Student-t: calls 2k and 2k+1 per packet, u1 = 1 - U
    n = 5.0
    st = orc.fill_studentt_f32(e.copy(), per, n).reshape(2, per // 8, 8)
    ww = w.reshape(2, per // 8, 2, 8)
    a = 1 - ur(ww[:, :, 0, :]).astype(np.float64)
    b = ur(ww[:, :, 1, :]).astype(np.float64)
    want = np.sqrt(n * (np.exp(np.log(a) * (-2 / n)) - 1)) * np.sin(2 * math.pi * b)
    np.testing.assert_allclose(st, want, rtol=3e-5, atol=3e-5)


def urd(w64):
    """bit_to_ur_double of u64 words (52 mantissa bits)."""
    return (w64 & np.uint64((1 << 52) - 1)).astype(np.float64) / 2.0 ** 52


@pytest.mark.parametrize("per", [1024, 1020])
def test_normal_f64_is_box_muller_of_the_stream(orc, per):
    # Packet4d: calls 2m, 2m+1 give u1, u2; outputs [r sin x4, r cos x4]; a
    # block of 4 mod 8 ends with a sine packet and consumes its two calls.
    e = orc.Engines(n=3, seed0=91)
    units = (per + 7) // 8
    got = orc.fill_normal_f64(e.copy(), per).reshape(3, per)
    raw = orc.fill_u64(e.copy(), 2 * units).reshape(3, units, 2, 4)
    u1 = 1.0 - urd(raw[:, :, 0, :])
    u2 = urd(raw[:, :, 1, :])
    r = np.sqrt(-2 * np.log(u1))
    want = np.stack([r * np.sin(2 * math.pi * u2), r * np.cos(2 * math.pi * u2)], 2)
    want = want.reshape(3, units * 8)[:, :per]
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-8 * (1 + np.abs(want)).max())
    # the engines advance by whole units
    e2 = e.copy()
    orc.fill_normal_f64(e2, per)
    e3 = e.copy()
    orc.fill_u64(e3, 2 * units)
    assert np.array_equal(orc.fill_u64(e2, 4), orc.fill_u64(e3, 4))
    # NormalGen: z * sd + mean, unfused, from the same z
    aff = orc.fill_normal_f64(e.copy(), per, 2.5, 0.25).reshape(3, per)
    assert np.array_equal(aff, got * 0.25 + 2.5)


def test_exponential_cauchy_f64_of_the_stream(orc):
    e = orc.Engines(n=2, seed0=17)
    per = 512
    u = urd(orc.fill_u64(e.copy(), per // 4)).reshape(2, per)
    ex = orc.fill_exponential_f64(e.copy(), per, 2.0).reshape(2, per)
    np.testing.assert_allclose(ex, -np.log(1 - u) / 2.0, rtol=0, atol=2e-9)
    assert np.all(np.signbit(ex[u == 0]))  # pnegate of +0
    ca = orc.fill_cauchy_f64(e.copy(), per, 1.0, 3.0).reshape(2, per)
    want = 1.0 + 3.0 * np.tan(math.pi * (u - 0.5))
    np.testing.assert_allclose(ca, want, rtol=1e-7, atol=1e-7)


@pytest.mark.parametrize("alpha", [1.0, 2.0, 7.0])
def test_gamma_integer_alpha_of_the_stream(orc, alpha):
    e = orc.Engines(n=2, seed0=9)
    per, beta = 256, f32(1.5)
    k = int(alpha)
    got = orc.fill_gamma_f32(e.copy(), per, alpha, float(beta)).reshape(2, per // 8, 8)
    w = full_words(orc, e, k * per // 8).reshape(2, per // 8, k, 8)
    if k == 1:  # beta * ExponentialGen: -(plog(1 - U) / 1)
        want = beta * -(orc.plog(f32(1) - ur(w[:, :, 0, :])) / f32(1))
    else:       # -(beta * plog(prod (U + eps/8))), multiplied in draw order
        y = ur(w[:, :, 0, :]) + EPS8
        for c in range(1, k):
            y = y * (ur(w[:, :, c, :]) + EPS8)
        want = -(beta * orc.plog(y))
    np.testing.assert_array_equal(got, want.astype(f32))


def _queue(W):
    """CompressMask<W> queue (arch/AVX/PacketFilter.h:156-177): a fully
    accepted packet bypasses; otherwise accepted lanes append in lane order
    and W of them are emitted once available."""
    rest = []

    def push(cands, acc):
        nonlocal rest
        if all(acc):
            return list(cands)
        rest = rest + [c for c, a in zip(cands, acc) if a]
        if len(rest) >= W:
            out, rest = rest[:W], rest[W:]
            return out
        return None
    return push


@pytest.mark.parametrize("alpha", [0.3, 0.8])
def test_gamma_gs_rejection_of_the_stream(orc, alpha):
    e = orc.Engines(n=2, seed0=31)
    per, beta = 64, f32(2.0)
    got = orc.fill_gamma_f32(e.copy(), per, alpha, float(beta)).reshape(2, per)
    a = f32(alpha)
    px = f32(math.e / (float(a) + math.e))
    w = full_words(orc, e, 600)
    for g in range(2):
        push, out, d = _queue(8), [], 0
        while len(out) < per:
            ux, vx, u3 = ur(w[g, d]), ur(w[g, d + 1]) + EPS8, ur(w[g, d + 2])
            d += 3
            lv = orc.plog(vx)
            x1 = orc.pexp((f32(1) / a) * lv)
            q1 = orc.pexp(-x1)
            x2 = f32(1) - orc.plog(vx)
            q2 = orc.pexp(orc.plog(x2) * (a - f32(1)))
            c = ux < px
            xx, qx = np.where(c, x1, x2), np.where(c, q1, q2)
            r = push(list(beta * xx), list(u3 < qx))
            if r is not None:
                out += r
        np.testing.assert_array_equal(got[g], np.array(out[:per], dtype=f32))


def test_gamma_tangent_rejection_of_the_stream(orc):
    e = orc.Engines(n=2, seed0=44)
    per, alpha, beta = 64, f32(3.7), f32(1.0)
    got = orc.fill_gamma_f32(e.copy(), per, float(alpha), float(beta)).reshape(2, per)
    sq = np.sqrt(f32(2) * alpha - f32(1))
    a1 = alpha - f32(1)
    w = full_words(orc, e, 400)
    for g in range(2):
        push, out, d = _queue(8), [], 0
        while len(out) < per:
            ys, yc = orc.psincos(f32(math.pi) * ur(w[g, d]))
            yx = ys / yc
            xx = sq * yx + a1
            ux = ur(w[g, d + 1])
            d += 2
            ub = (yx * yx + f32(1)) * orc.pexp(a1 * orc.plog(xx / a1) - yx * sq)
            acc = (f32(0) < xx) & (ux <= ub)
            r = push(list(beta * xx), list(acc))
            if r is not None:
                out += r
        np.testing.assert_array_equal(got[g], np.array(out[:per], dtype=f32))


@pytest.mark.parametrize("mean", [0.05, 3.0, 11.5])
def test_poisson_knuth_of_the_stream(orc, mean):
    e = orc.Engines(n=2, seed0=101)
    per = 128
    got = orc.fill_poisson_i32(e.copy(), per, mean).reshape(2, per // 4, 4)
    h = half_words(orc, e, 4000)
    thr = f32(math.exp(-mean))
    for g in range(2):
        s = 0
        for k in range(per // 4):
            val = np.ones(4, f32)
            res = np.zeros(4, np.int32)
            while True:
                val = val * ur(h[g, s])
                s += 1
                c = thr < val
                if not c.any():
                    break
                res += c
            np.testing.assert_array_equal(got[g, k], res)


def test_poisson_tangent_rejection_of_the_stream(orc):
    e = orc.Engines(n=2, seed0=7)
    per, mean = 64, 16.0
    got = orc.fill_poisson_i32(e.copy(), per, mean).reshape(2, per)
    psq, pm = f32(math.sqrt(2 * mean)), f32(mean)
    plm, pg1 = f32(math.log(mean)), f32(mean * math.log(mean) - math.lgamma(mean + 1))
    h = half_words(orc, e, 400)
    for g in range(2):
        push, out, s = _queue(4), [], 0
        while len(out) < per:
            ps, pc = orc.psincos(f32(math.pi) * ur(h[g, s]))
            yx = ps / pc
            fres = np.trunc(psq * yx + pm).astype(f32)
            p1 = (yx * yx + f32(1)) * f32(0.9)
            p2 = orc.pexp(fres * plm - orc.plgamma(fres + f32(1)) - pg1)
            acc = (f32(0) <= fres) & (ur(h[g, s + 1]) <= p1 * p2)
            s += 2
            r = push(list(fres), list(acc))
            if r is not None:
                out += r
        np.testing.assert_array_equal(got[g], np.array(out[:per], dtype=f32).astype(np.int32))


@pytest.mark.parametrize("n,p", [(7, 0.3), (24, 0.81)])
def test_binomial_small_n_of_the_stream(orc, n, p):
    e = orc.Engines(n=2, seed0=3)
    per = 64
    got = orc.fill_binomial_i32(e.copy(), per, n, p).reshape(2, per // 4, 4)
    h = half_words(orc, e, n * per // 4).reshape(2, per // 4, n, 4)
    # res = n + sum_i (p <= U_i ? -1 : 0), one half draw per trial
    want = n - (f32(p) <= ur(h)).sum(axis=2)
    np.testing.assert_array_equal(got, want)


def test_binomial_tangent_rejection_of_the_stream(orc):
    e = orc.Engines(n=2, seed0=12)
    per, n, p = 64, 100, 0.75
    got = orc.fill_binomial_i32(e.copy(), per, n, p).reshape(2, per)
    sp = min(p, 1 - p)
    mean = n * sp
    ptr, pm = f32(n), f32(mean)
    psv = f32(math.sqrt(2 * mean * (1 - sp)))
    plp, plq, pg1 = f32(math.log(sp)), f32(math.log(1 - sp)), f32(math.lgamma(n + 1))
    h = half_words(orc, e, 600)
    for g in range(2):
        push, out, s = _queue(4), [], 0
        while len(out) < per:
            ps, pc = orc.psincos(f32(math.pi) * ur(h[g, s]))
            ys = ps / pc
            fres = np.trunc(psv * ys + pm).astype(f32)
            p1 = (f32(1.2) * psv) * (f32(1) + ys * ys)
            p2 = orc.pexp(((pg1 - orc.plgamma(fres + f32(1))) - orc.plgamma((ptr + f32(1)) - fres))
                          + fres * plp + (ptr - fres) * plq)
            acc = (f32(0) <= fres) & (fres <= ptr) & (ur(h[g, s + 1]) <= p1 * p2)
            s += 2
            r = push(list(fres), list(acc))
            if r is not None:
                out += r
        res = np.array(out[:per], dtype=f32).astype(np.int32)
        np.testing.assert_array_equal(got[g], n - res)  # p != small_p: n - k


def test_lognormal_weibull_extreme_value_of_the_stream(orc):
    e = orc.Engines(n=2, seed0=21)
    per = 512
    w = full_words(orc, e, per // 8)
    u = ur(w).astype(np.float64).reshape(2, per)
    wb = orc.fill_weibull_f32(e.copy(), per, 1.7, 2.5).reshape(2, per)
    np.testing.assert_allclose(wb, 2.5 * (-np.log(1 - u)) ** (1 / 1.7), rtol=2e-5, atol=1e-6)
    ev = orc.fill_extreme_value_f32(e.copy(), per, 1.0, 3.0).reshape(2, per)
    unz = u + 2.0 ** -26
    np.testing.assert_allclose(ev, 1.0 - 3.0 * np.log(-np.log(unz)), rtol=2e-5, atol=2e-5)
    # lognormal = exp of the NormalGen fill, packet for packet
    ln = orc.fill_lognormal_f32(e.copy(), per, 0.5, 0.75)
    nz = orc.fill_normal_f32(e.copy(), per, 0.5, 0.75)
    np.testing.assert_allclose(ln, np.exp(nz.astype(np.float64)), rtol=3e-7 * 4)


def test_bernoulli_balanced_randbits_of_the_stream(orc):
    """Raw-bit generators against the pinned u64 stream: 31 bits against
    uint32(p * 2^31) as signed int32 (so p = 1 gives zeros on the packet path),
    zero_to_one = int(w & 0x7FFFFFFF) / 2^31 in float, and the half() words."""
    e = orc.Engines(n=2, seed0=5)
    per = 4096
    w = full_words(orc, e, per // 8).reshape(2, per)
    r31 = (w & 0x7FFFFFFF).astype(np.int64)
    for p in (0.3, 0.5, 1.0, 0.0):
        got = orc.fill_bernoulli_f32(e.copy(), per, p).reshape(2, per)
        t = np.int64(np.uint32(p * 2.0 ** 31).astype(np.int32))
        np.testing.assert_array_equal(got, (r31 < t).astype(f32))
    assert not orc.fill_bernoulli_f32(e.copy(), per, 1.0).any()
    assert abs(orc.fill_bernoulli_f32(e.copy(), 1 << 16, 0.3).mean() - 0.3) < 0.01
    z = (r31.astype(f32) / f32(2.0 ** 31)).astype(f32)
    np.testing.assert_array_equal(orc.fill_balanced_f32(e.copy(), per).reshape(2, per), z * f32(2) - f32(1))
    ab = orc.fill_balanced_f32(e.copy(), per, -3.0, 5.5).reshape(2, per)
    np.testing.assert_array_equal(ab, z * (f32(5.5) - f32(-3.0)) + f32(-3.0))
    big = orc.fill_balanced_f32(e.copy(), 1 << 16, -3.0, 5.5)
    assert big.min() >= -3.0 and big.max() <= 5.5
    assert abs(big.mean() - 1.25) < 0.05 and abs(big.var() - 8.5 ** 2 / 12) < 0.1
    # int32 forms through half()
    h = half_words(orc, e, per // 4).reshape(2, per)
    np.testing.assert_array_equal(orc.fill_randbits_i32(e.copy(), per).reshape(2, per), h.view(np.int32))
    bi = orc.fill_bernoulli_i32(e.copy(), per, 0.3).reshape(2, per)
    np.testing.assert_array_equal(bi, ((h & 0x7FFFFFFF).astype(np.int64) < np.int64(int(0.3 * 2 ** 31))).astype(np.int32))


def test_geometric_of_the_stream(orc):
    """trunc(log(1 - U) / log(1 - p)) from half() words: equal to a float64
    evaluation except where the quotient lies within float rounding of an
    integer; the distribution is Geometric(p) on {0, 1, ...}."""
    e = orc.Engines(n=2, seed0=9)
    per = 1 << 15
    p = 0.3
    got = orc.fill_geometric_i32(e.copy(), per, p).reshape(2, per)
    h = half_words(orc, e, per // 4).reshape(2, per)
    u = ur(h).astype(np.float64)
    ref = np.trunc(np.log(1 - u) * np.float64(np.float32(1 / np.log(1 - p)))).astype(np.int64)
    diff = got - ref
    assert np.abs(diff).max() <= 1 and (diff != 0).mean() < 1e-3
    x = got.ravel()
    assert abs(x.mean() - (1 - p) / p) < 0.05
    for k in range(4):
        assert abs((x == k).mean() - (1 - p) ** k * p) < 0.01
    # p = 0: 1/log(1) = inf -> every product is -inf or NaN -> INT32_MIN (cvttps)
    assert (orc.fill_geometric_i32(e.copy(), 64, 0.0) == np.iinfo(np.int32).min).all()


@pytest.mark.parametrize("kind,scalar", [
    ("uniform", lambda o, e, n, a, b: o.fill_uniform_f32(e, n, a, b)),
    ("exponential", lambda o, e, n, a, b: o.fill_exponential_f32(e, n, a)),
    ("cauchy", lambda o, e, n, a, b: o.fill_cauchy_f32(e, n, a, b)),
    ("weibull", lambda o, e, n, a, b: o.fill_weibull_f32(e, n, a, b)),
    ("extreme_value", lambda o, e, n, a, b: o.fill_extreme_value_f32(e, n, a, b)),
    ("studentt", lambda o, e, n, a, b: o.fill_studentt_f32(e, n, a)),
    ("lognormal", lambda o, e, n, a, b: o.fill_lognormal_f32(e, n, a, b)),
])
def test_vop_with_constant_parameters_is_the_scalar_fill(orc, kind, scalar):
    """A vectorize-over-params fill whose parameters are all equal is the
    scalar generator's fill (the per-element formulas reduce to it), for two
    parameter sets; with varying parameters, element i depends on its own
    parameters only (checked by changing one block)."""
    e = orc.Engines(n=3, seed0=33)
    per = 1024
    for a, b in ((1.7, 2.5), (0.6, 4.0)):
        va = np.full(3 * per, a, np.float32)
        vb = np.full(3 * per, b, np.float32)
        np.testing.assert_array_equal(orc.fill_vop_f32(e.copy(), kind, va, vb),
                                      scalar(orc, e.copy(), per, a, b))
    rng = np.random.default_rng(1)
    va = rng.uniform(0.5, 3, 3 * per).astype(np.float32)
    vb = rng.uniform(0.5, 3, 3 * per).astype(np.float32)
    x = orc.fill_vop_f32(e.copy(), kind, va, vb)
    va2 = va.copy()
    va2[per:2 * per] += 1
    y = orc.fill_vop_f32(e.copy(), kind, va2, vb)
    np.testing.assert_array_equal(x[:per], y[:per])
    np.testing.assert_array_equal(x[2 * per:], y[2 * per:])


def test_uniform_and_balanced_vop_are_affine_maps_of_the_stream(orc):
    e = orc.Engines(n=2, seed0=3)
    per = 2048
    rng = np.random.default_rng(2)
    a = rng.uniform(-5, 5, 2 * per).astype(np.float32)
    b = (a + rng.uniform(0.1, 5, 2 * per)).astype(np.float32)
    w = full_words(orc, e, per // 8).reshape(2 * per)
    u = ur(w)
    np.testing.assert_array_equal(orc.fill_vop_f32(e.copy(), "uniform", a, b), u * (b - a) + a)
    z = ((w & 0x7FFFFFFF).astype(np.int64).astype(f32) / f32(2.0 ** 31)).astype(f32)
    np.testing.assert_array_equal(orc.fill_balanced_vop_f32(e.copy(), a, b), z * (b - a) + a)
    p = rng.uniform(0, 1, 2 * per).astype(np.float32)
    t = (p * f32(2.0 ** 31)).astype(np.float64)
    thr = np.where(t >= 2.0 ** 31, -(2 ** 31), np.trunc(t)).astype(np.int64)
    np.testing.assert_array_equal(orc.fill_bernoulli_vop_f32(e.copy(), p),
                                  ((w & 0x7FFFFFFF).astype(np.int64) < thr).astype(f32))


def _uniform_int_reference(words, lo, hi, n_packets):
    """UniformIntGen<int32_t>::packetOp written out on the half() word stream
    (4 words per draw) with an explicit CompressMask<16> FIFO."""
    pdiff = hi - lo
    bits = int(np.ceil(np.log2(pdiff + 1)))
    mask = (1 << bits) - 1
    it = iter(words.reshape(-1, 4))
    out, rest = [], []
    for _ in range(n_packets):
        rx = [int(x) for x in next(it)]
        if pdiff == mask:
            out += [((x & mask) + lo) for x in rx]
            continue
        bitcnt = bits
        while True:
            cands = [x & mask for x in rx]
            acc = [c < pdiff + 1 for c in cands]
            if all(acc):
                out += [c + lo for c in cands]
                break
            rest += [c for c, a in zip(cands, acc) if a]
            if len(rest) >= 4:
                out += [c + lo for c in rest[:4]]
                rest = rest[4:]
                break
            if bitcnt + bits < 32:
                rx = [x >> bits for x in rx]
                bitcnt += bits
            else:
                rx = [int(x) for x in next(it)]
                bitcnt = bits
    return np.array(out, dtype=np.int64)


@pytest.mark.parametrize("lo,hi", [(0, 9), (-3, 4), (0, 2), (10, 1000), (0, 7), (5, 5), (-100, 2 ** 30),
                                   (0, 2 ** 31 - 1)])
def test_uniform_int_of_the_stream(orc, lo, hi):
    e = orc.Engines(n=1, seed0=17)
    per = 2048
    got = orc.fill_uniform_int_i32(e.copy(), per, lo, hi).astype(np.int64)
    if hi == lo:
        assert (got == lo).all()
        f = e.copy()
        orc.fill_uniform_int_i32(f, per, lo, hi)
        np.testing.assert_array_equal(f.raw, e.raw)  # a constant range draws nothing
        return
    words = half_words(orc, e, 8 * per)[0]
    np.testing.assert_array_equal(got, _uniform_int_reference(words, lo, hi, per // 4))
    big = orc.fill_uniform_int_i32(e.copy(), 1 << 18, lo, hi).astype(np.int64)
    assert big.min() >= lo and big.max() <= hi
    if hi - lo < 64:  # every value with probability 1/(hi-lo+1)
        counts = np.bincount(big - lo, minlength=hi - lo + 1)
        exp = big.size / (hi - lo + 1)
        assert ((counts - exp) ** 2 / exp).sum() < 3 * (hi - lo) + 30
    else:
        assert abs(big.mean() - (lo + hi) / 2) < 0.01 * (hi - lo)


@pytest.mark.parametrize("K", [3, 15, 16, 37, 1000])
def test_discrete_f_of_the_stream(orc, K):
    """discreteF: K < 16 counts a half draw's uniform against the float CDF
    (the category is the first c_i above U); K >= 16 takes 2^b-masked bits
    from one half draw and U from the next against the float alias table;
    frequencies match the weights."""
    rng = np.random.default_rng(K)
    w = rng.uniform(0.1, 1, K)
    e = orc.Engines(n=1, seed0=23)
    per = 4096
    got = orc.fill_discrete_f_i32(e.copy(), per, w)
    h = half_words(orc, e, 2 * per // 4)[0]
    if K < 16:
        c = np.cumsum(w)
        cdf = (c / c[-1]).astype(f32)
        u = ur(h[:per // 4]).reshape(-1)
        exp = np.array([K - int((x < cdf).sum()) for x in u])
        np.testing.assert_array_equal(got, exp)
    else:
        assert got.min() >= 0 and got.max() < K
    big = orc.fill_discrete_f_i32(e.copy(), 1 << 18, w)
    freq = np.bincount(big, minlength=K) / big.size
    assert np.abs(freq - w / w.sum()).max() < 6 * np.sqrt(w.max() / w.sum() / big.size) + 1e-3


@pytest.mark.parametrize("a,b", [(0.4, 0.7), (0.5, 0.5), (2.0, 3.0), (0.5, 4.5), (1.0, 1.0), (3.3, 0.2)])
def test_beta_and_fisher_f_of_the_stream(orc, a, b):
    """BetaGen: Johnk (a, b < 1) accepted candidates p1/(p1+p2) are the
    stream's own draws (checked by replaying the queue), and the gamma-ratio
    regime equals GammaGen(a) and GammaGen(b) packets drawn in turn; moments
    match Beta(a, b); FisherF is its (b/a) x/(1-x) map."""
    e = orc.Engines(n=1, seed0=41)
    per = 1024
    got = orc.fill_beta_f32(e.copy(), per, a, b)
    if a < 1 and b < 1:
        w = full_words(orc, e, 8 * per // 8)[0]  # plenty of draws
        u = ur(w)
        outs, rest, k = [], [], 0
        while len(outs) < per:
            u1, u2 = u[k], u[k + 1]
            k += 2
            p1 = orc.pexp(orc.plog(u1) * f32(1 / f32(a)))
            p2 = orc.pexp(orc.plog(u2) * f32(1 / f32(b)))
            x = (p1 + p2).astype(f32)
            c = (p1 / x).astype(f32)
            acc = x <= 1
            if acc.all():
                outs += list(c)
                continue
            rest += list(c[acc])
            if len(rest) >= 8:
                outs += rest[:8]
                rest = rest[8:]
        np.testing.assert_array_equal(got, np.array(outs[:per], dtype=f32))
    else:
        # one GammaGen(a) packet, then one GammaGen(b) packet, per output packet
        ref, h = [], e.copy()
        g1 = orc.Generators(9, a, 1.0, 1)
        g2 = orc.Generators(9, b, 1.0, 1)
        for _ in range(per // 8):
            p1 = g1.fill(h, 8)
            p2 = g2.fill(h, 8)
            ref += list((p1 / (p1 + p2)).astype(f32))
        np.testing.assert_array_equal(got, np.array(ref, dtype=f32))
    big = orc.fill_beta_f32(e.copy(), 1 << 16, a, b).astype(np.float64)
    mean = a / (a + b)
    var = a * b / ((a + b) ** 2 * (a + b + 1))
    assert abs(big.mean() - mean) < 0.01 and abs(big.var() - var) < 0.01
    fz = orc.fill_fisher_f_f32(e.copy(), per, 2 * a, 2 * b)
    bz = orc.fill_beta_f32(e.copy(), per, a, b)
    r = f32(b) / f32(a)
    with np.errstate(divide="ignore"):  # x = 1 maps to +inf on both sides
        np.testing.assert_array_equal(fz, (r * bz) / (f32(1) - bz))


@pytest.mark.parametrize("trials,p", [(1, 0.5), (3, 0.7), (10, 0.5), (25, 0.3), (50, 0.05), (19, 0.9)])
def test_negative_binomial_moments(orc, trials, p):
    """NegativeBinomialGen = Poisson(Gamma(trials, (1-p)/p)): mean
    trials (1-p)/p and variance trials (1-p)/p^2 within sampling error."""
    e = orc.Engines(n=4, seed0=3)
    x = orc.fill_negative_binomial_i32(e, 1 << 15, trials, p).astype(np.float64)
    m = trials * (1 - p) / p
    v = trials * (1 - p) / p ** 2
    assert x.min() >= 0
    assert abs(x.mean() - m) < 5 * np.sqrt(v / x.size) + 1e-3
    assert abs(x.var() / v - 1) < 0.05


# ------------------------------------------------------------ DiscreteGen ---
def _discrete_cat(words, w, prob=None, alias=None):
    """packetOp category of 32-bit words (Discrete.h:446-471): K < 16 counts
    the CDF thresholds above the 31-bit value (signed compares); otherwise the
    alias test (word >> 1) < prob[word & bitmask] picks the slot or its alias."""
    words = words.astype(np.uint32)
    K = len(w)
    if K < 16:
        c = np.cumsum(np.asarray(w, np.float64))
        cdf = (c / c[-1] * 0x80000000).astype(np.uint64).astype(np.uint32).view(np.int32)
        rx = (words & np.uint32(0x7FFFFFFF)).view(np.int32)
        ret = np.full(words.shape, K - 1, np.int32)
        for i in range(K - 1):
            ret -= (rx < cdf[i]).astype(np.int32)
        return ret
    albit = (words & np.uint32(prob.size - 1)).astype(np.int64)
    c = (words >> np.uint32(1)).view(np.int32) < prob[albit]
    return np.where(c, albit, alias[albit]).astype(np.int32)


@pytest.mark.parametrize("K,per", [(5, 1024), (15, 1020), (16, 1024), (100, 1028), (1000, 512)])
def test_discrete_of_the_stream(orc, K, per):
    # AVX2: one full draw gives a Packet8i; packet 2k is its low half, packet
    # 2k+1 the cached high half; a fill of 4 mod 8 ends on a low half and
    # drops the cached one with the generator.
    rng = np.random.default_rng(K)
    w = rng.uniform(0, 1, K)
    w[::7] = 0.0
    e = orc.Engines(n=2, seed0=1 + K)
    got = orc.fill_discrete_i32(e.copy(), per, w).reshape(2, per)
    draws = (per + 7) // 8
    words = full_words(orc, e, draws).reshape(2, draws * 8)[:, :per]
    prob = alias = None
    if K >= 16:
        prob, alias = orc.alias_table(w)
    np.testing.assert_array_equal(got, _discrete_cat(words, w, prob, alias))
    e2 = e.copy()
    orc.fill_discrete_i32(e2, per, w)
    e3 = e.copy()
    orc.fill_u64(e3, draws)
    np.testing.assert_array_equal(e2.raw, e3.raw)


# ------------------------------------------------------- BinomialVGen<int> ---
def _binomial_v_packet(orc, h, s, n, p):
    """BinomialVGen<int32_t>::packetOp (Discrete.h:1124-1190) on half draws
    h[s:], four lanes; returns (values, half draws consumed)."""
    return _binomial_v_lanes(orc, lambda q: ur(h[q]), s, n, p)


def _binomial_v_lanes(orc, u, s, n, p):
    """The BinomialVGen packetOp body (Discrete.h:1124-1190 for int32 lanes,
    1257-1327 for float lanes, the same steps): u(q) gives the uniforms of
    draw q for all lanes. Returns (values, draws consumed)."""
    ps = np.minimum(p, f32(1) - p)
    tr = n.astype(f32)
    pmean = tr * ps
    pg1 = orc.plgamma(tr + f32(1))
    psq = np.sqrt(f32(2) * pmean * (f32(1) - ps)).astype(f32)
    plp = orc.plog(ps)
    plq = orc.plog(f32(1) - ps)
    valid = pmean < f32(1)
    L = n.size
    res = np.zeros(L, f32)
    s0 = s
    if valid.any():
        nem = orc.pexp(-pmean)
        val = np.ones(L, f32)
        cnt = np.zeros(L, np.int32)
        while True:
            val = (val * u(s)).astype(f32)
            s += 1
            c = (nem < val) & valid
            if not c.any():
                break
            cnt += c
        res = cnt.astype(f32)
    if not valid.all():
        while True:
            sn, cs = orc.psincos(f32(math.pi) * u(s))
            ys = (sn / cs).astype(f32)
            fres = np.trunc(psq * ys + pmean).astype(f32)
            p1 = (f32(1.2) * psq) * (f32(1) + ys * ys)
            with np.errstate(invalid="ignore"):  # p = 0 or 1 lanes: resolved by the Knuth loop
                p2 = orc.pexp((((pg1 - orc.plgamma(fres + f32(1))) - orc.plgamma((tr + f32(1)) - fres))
                               + fres * plp) + (tr - fres) * plq)
            c1 = (f32(0) <= fres) & (fres <= tr)
            c2 = u(s + 1) <= (p1 * p2).astype(f32)
            s += 2
            fv = c1 & c2
            res = np.where(fv & ~valid, fres, res)
            valid = valid | fv
            if valid.all():
                break
    k = np.trunc(res).astype(np.int32)
    return np.where(p == ps, k, n - k), s - s0


def test_binomial_vop_of_the_stream(orc):
    rng = np.random.default_rng(31)
    per = 256
    n = rng.integers(1, 1001, 2 * per).astype(np.int32)
    p = rng.uniform(0, 1, 2 * per).astype(f32)
    n[::5] = rng.integers(1, 30, n[::5].size)       # small means: the Knuth lockstep
    p[::7] = f32(0.003)
    n[::11] = 0
    p[::13] = f32(1.0)
    e = orc.Engines(n=2, seed0=2024)
    got = orc.fill_binomial_vop_i32(e.copy(), n, p).reshape(2, per)
    h = half_words(orc, e, 40 * per)
    for g in range(2):
        s = 0
        for k in range(per // 4):
            sl = slice(g * per + 4 * k, g * per + 4 * k + 4)
            v, used = _binomial_v_packet(orc, h[g], s, n[sl], p[sl])
            s += used
            np.testing.assert_array_equal(got[g, 4 * k:4 * k + 4], v, err_msg=f"engine {g} packet {k}")


def test_binomial_vop_float_of_the_stream(orc):
    """BinomialVGen<float> (Discrete.h:1251-1330): the same packet steps on
    Packet8f lanes of full draws, per-element trials and p."""
    rng = np.random.default_rng(37)
    per = 256
    n = rng.integers(1, 1001, 2 * per).astype(np.int32)
    p = rng.uniform(0, 1, 2 * per).astype(f32)
    n[::5] = rng.integers(1, 30, n[::5].size)
    p[::7] = f32(0.003)
    n[::11] = 0
    p[::13] = f32(1.0)
    e = orc.Engines(n=2, seed0=4048)
    got = orc.fill_binomial_vopf_i32(e.copy(), n, p).reshape(2, per)
    words = full_words(orc, e, 40 * per)
    for g in range(2):
        s = 0
        for k in range(per // 8):
            sl = slice(g * per + 8 * k, g * per + 8 * k + 8)
            v, used = _binomial_v_lanes(orc, lambda q: ur(words[g][q]), s, n[sl], p[sl])
            s += used
            np.testing.assert_array_equal(got[g, 8 * k:8 * k + 8], v, err_msg=f"engine {g} packet {k}")


# ----------------------------------------------------------- MultinomialGen ---
def _multinomial_replay(orc, e, trials, w, samples):
    """MultinomialGen<int32_t>::generate(urng, samples) (Multinomial.h:60-133)
    on the raw full-draw stream of engine e; returns (K, samples)."""
    K = len(w)
    w = np.asarray(w, np.float64)
    probs = w / w.sum()
    words = full_words(orc, e, 200000)[0]  # (draws, 8) u32
    pos = [0]
    out = np.zeros((K, samples), np.int32)
    if trials < max(30 * (K - 1), 100):
        prob = alias = None
        if K >= 16:
            prob, alias = orc.alias_table(w)
        cache = [None]

        def packet4():  # AVX2: a Packet8i per draw, the high half cached
            if cache[0] is not None:
                c, cache[0] = cache[0], None
                return c
            cats = _discrete_cat(words[pos[0]], w, prob, alias)
            pos[0] += 1
            cache[0] = cats[4:]
            return cats[:4]

        def scalar():  # operator() (Discrete.h:428-441): word 0 of a fresh draw
            rx = np.int32(words[pos[0], 0].view(np.int32))
            pos[0] += 1
            if K < 16:
                c = np.cumsum(w)
                cdf = (c / c[-1] * 0x80000000).astype(np.uint64)
                r = int(rx) & 0x7FFFFFFF
                return int(np.searchsorted(cdf[:-1], r, side="left"))
            albit = int(rx) & (prob.size - 1)
            alx = (int(rx) >> 1) & 0xFFFFFFFF  # arithmetic shift, then as uint32
            return albit if alx < int(prob[albit]) else int(alias[albit])

        for s in range(samples):
            i = 0
            while i < (trials & ~15):
                buf = np.concatenate([packet4() for _ in range(4)])
                np.add.at(out[:, s], buf, 1)
                i += 16
            r = trials - i
            buf, j = [], 0
            while j + 4 <= r:
                buf += list(packet4())
                j += 4
            while j < r:
                buf.append(scalar())
                j += 1
            np.add.at(out[:, s], np.asarray(buf, np.int64), 1)
        return out
    r = np.full(samples, trials, np.int32)
    rest_p = 1 - probs[0]
    for row in range(K - 1):
        p = f32(probs[0] if row == 0 else probs[row] / rest_p)
        t = np.zeros(samples, np.int32)
        for q in range(0, samples, 8):
            t[q:q + 8], used = _binomial_v_lanes(orc, lambda d: ur(words[d]), pos[0], r[q:q + 8],
                                                 np.full(8, p, f32))
            pos[0] += used
        r -= t
        out[row] = t
        if row > 0:
            rest_p -= probs[row]
    out[K - 1] = r
    return out


@pytest.mark.parametrize("trials,K,samples", [(20, 4, 8), (99, 10, 16), (7, 3, 8), (50, 20, 8),
                                              (200, 40, 8), (1000, 4, 16), (5000, 12, 24), (3001, 3, 8)])
def test_multinomial_of_the_stream(orc, trials, K, samples):
    rng = np.random.default_rng(trials + K)
    w = rng.uniform(0.05, 1, K)
    e = orc.Engines(n=1, seed0=trials)
    got = orc.multinomial(e.copy(), trials, [w], samples)[0]
    np.testing.assert_array_equal(got, _multinomial_replay(orc, e, trials, w, samples))


# ---------------------------------------------------- NegativeBinomialGen ---
def _cvtt(x):
    """cvttps2dq: truncation; NaN and out-of-range give INT32_MIN."""
    x = np.asarray(x, np.float64)
    ok = np.isfinite(x) & (x < 2.0 ** 31) & (x >= -2.0 ** 31)
    return np.where(ok, np.trunc(np.where(ok, x, 0)), -2.0 ** 31).astype(np.int64).astype(np.int32)


@pytest.mark.parametrize("trials,p", [(1, 0.5), (3, 0.7), (10, 0.5), (19, 0.9), (25, 0.3), (50, 0.05)])
def test_negative_binomial_of_the_stream(orc, trials, p):
    """NegativeBinomialGen::packetOp (GammaPoisson.h:58-105) at Packet4i on
    half draws: the GammaGen packet (NormalExp.h:527-602: exponential for
    alpha = 1, a product of nonzero uniforms for integer alpha < 20, else the
    tangent step through its CompressMask queue), then Knuth on all lanes if
    no mean exceeds 12, else Poisson tangent candidates appended to the SAME
    queue (the gamma's rest buffer) and emitted through cvttps."""
    per = 256
    e = orc.Engines(n=1, seed0=trials * 7 + 1)
    got = orc.fill_negative_binomial_i32(e.copy(), per, trials, p)
    h = half_words(orc, e, 60000)[0]
    alpha, beta = f32(trials), f32((1 - p) / p)
    push = _queue(4)  # the gamma's queue, shared with the Poisson tangent stage
    s = 0
    out = []
    eps8 = f32(2.0 ** -26)
    while len(out) < per:
        # GammaGen<float>::packetOp<Packet4f>
        if alpha == f32(1):
            mean = beta * -(orc.plog(f32(1) - ur(h[s])) / f32(1))
            s += 1
        elif trials < 20:
            yx = ur(h[s]) + eps8
            s += 1
            for _ in range(trials - 1):
                yx = (yx * (ur(h[s]) + eps8)).astype(f32)
                s += 1
            mean = -(beta * orc.plog(yx))
        else:
            a1 = alpha - f32(1)
            sq = np.sqrt(f32(2) * alpha - f32(1)).astype(f32)
            while True:
                ys, yc = orc.psincos(f32(math.pi) * ur(h[s]))
                yx = (ys / yc).astype(f32)
                xx = sq * yx + a1
                c = f32(0) < xx
                ux = ur(h[s + 1])
                s += 2
                with np.errstate(invalid="ignore", divide="ignore"):
                    ub = (yx * yx + f32(1)) * orc.pexp(a1 * orc.plog(xx / a1) - yx * sq)
                r = push(list(beta * xx), list(c & (ux <= ub)))
                if r is not None:
                    mean = np.array(r, f32)
                    break
        mean = mean.astype(f32)
        # the Poisson stage
        nem = orc.pexp(-mean)
        if not (f32(12) < mean).any():
            val = np.ones(4, f32)
            res = np.zeros(4, np.int32)
            while True:
                val = (val * ur(h[s])).astype(f32)
                s += 1
                c = nem < val
                if not c.any():
                    break
                res += c
            out += list(res)
        else:
            psq = np.sqrt(f32(2) * mean).astype(f32)
            plm = orc.plog(mean)
            pg1 = mean * plm - orc.plgamma(mean + f32(1))
            while True:
                sn, cs = orc.psincos(f32(math.pi) * ur(h[s]))
                yx = (sn / cs).astype(f32)
                fres = np.trunc(psq * yx + mean).astype(f32)
                p1 = (yx * yx + f32(1)) * f32(0.9)
                with np.errstate(invalid="ignore", over="ignore"):
                    p2 = orc.pexp((fres * plm - orc.plgamma(fres + f32(1))) - pg1)
                acc = (f32(0) <= fres) & (ur(h[s + 1]) <= p1 * p2)
                s += 2
                r = push(list(fres), list(acc))
                if r is not None:
                    out += list(_cvtt(np.array(r, f32)))
                    break
    np.testing.assert_array_equal(got, np.array(out[:per], np.int32))


# ------------------------------------------------- gamma / Poisson VoP ---
# Build-defined (the reference has no GammaVGen/PoissonVGen): BinomialVGen's
# lockstep design with per-lane parameters; each lane keeps its first
# accepted candidate and a packet consumes the iterations of its slowest lane.

def test_gamma_vop_of_the_stream(orc):
    rng = np.random.default_rng(77)
    per = 128
    alpha = np.exp(rng.uniform(np.log(0.1), np.log(50), 2 * per)).astype(f32)
    alpha[::9] = f32(1.0)
    beta = rng.uniform(0.5, 2, 2 * per).astype(f32)
    e = orc.Engines(n=2, seed0=5)
    got = orc.fill_gamma_vop_f32(e.copy(), alpha, beta).reshape(2, per)
    w = full_words(orc, e, 6000)
    eps8 = f32(2.0 ** -26)
    for g in range(2):
        d = 0
        for k in range(per // 8):
            a = alpha[g * per + 8 * k: g * per + 8 * k + 8]
            b = beta[g * per + 8 * k: g * per + 8 * k + 8]
            px = f32(math.e) / (a + f32(math.e))
            inv = f32(1) / a
            am1 = a - f32(1)
            with np.errstate(invalid="ignore"):  # alpha < 0.5: unused (GS lanes)
                sq = np.sqrt(f32(2) * a - f32(1)).astype(f32)
            res = np.zeros(8, f32)
            done = np.zeros(8, bool)
            while not done.all():
                u1, u2, u3 = ur(w[g, d]), ur(w[g, d + 1]), ur(w[g, d + 2])
                d += 3
                with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
                    lv = orc.plog(u2 + eps8)                    # GS lanes (alpha <= 1)
                    x1 = orc.pexp(inv * lv)
                    q1 = orc.pexp(-x1)
                    x2 = f32(1) - lv
                    q2 = orc.pexp(orc.plog(x2) * am1)
                    c = u1 < px
                    xg = np.where(c, x1, x2)
                    ag = u3 < np.where(c, q1, q2)
                    ys, yc = orc.psincos(f32(math.pi) * u1)     # tangent lanes
                    yx = (ys / yc).astype(f32)
                    xt = sq * yx + am1
                    ub = (yx * yx + f32(1)) * orc.pexp(am1 * orc.plog(xt / am1) - yx * sq)
                    at = (f32(0) < xt) & (u2 <= ub)
                gs = a <= f32(1)
                x = np.where(gs, xg, xt)
                acc = np.where(gs, ag, at)
                new = acc & ~done
                res = np.where(new, b * x, res)
                done |= acc
            np.testing.assert_array_equal(got[g, 8 * k:8 * k + 8], res, err_msg=f"engine {g} packet {k}")


def test_poisson_vop_of_the_stream(orc):
    rng = np.random.default_rng(78)
    per = 256
    mean = np.exp(rng.uniform(np.log(0.01), np.log(1e4), 2 * per)).astype(f32)
    mean[::11] = f32(12.0)
    mean[::13] = f32(11.99)
    e = orc.Engines(n=2, seed0=6)
    got = orc.fill_poisson_vop_i32(e.copy(), mean).reshape(2, per)
    h = half_words(orc, e, 20000)
    for g in range(2):
        s = 0
        for k in range(per // 4):
            m = mean[g * per + 4 * k: g * per + 4 * k + 4]
            small = m < f32(12)
            nem = orc.pexp(-m)
            psq = np.sqrt(f32(2) * m).astype(f32)
            plm = orc.plog(m)
            pg1 = m * plm - orc.plgamma(m + f32(1))
            res = np.zeros(4, f32)
            if small.any():
                val = np.ones(4, f32)
                cnt = np.zeros(4, np.int32)
                while True:
                    val = (val * ur(h[g, s])).astype(f32)
                    s += 1
                    c = (nem < val) & small
                    if not c.any():
                        break
                    cnt += c
                res = cnt.astype(f32)
            if not small.all():
                done = small.copy()
                while not done.all():
                    sn, cs = orc.psincos(f32(math.pi) * ur(h[g, s]))
                    yx = (sn / cs).astype(f32)
                    fres = np.trunc(psq * yx + m).astype(f32)
                    p1 = (yx * yx + f32(1)) * f32(0.9)
                    with np.errstate(invalid="ignore", over="ignore"):
                        p2 = orc.pexp((fres * plm - orc.plgamma(fres + f32(1))) - pg1)
                    acc = (f32(0) <= fres) & (ur(h[g, s + 1]) <= p1 * p2)
                    s += 2
                    res = np.where(acc & ~done, fres, res)
                    done |= acc
            np.testing.assert_array_equal(got[g, 4 * k:4 * k + 4], res.astype(np.int32),
                                          err_msg=f"engine {g} packet {k}")
