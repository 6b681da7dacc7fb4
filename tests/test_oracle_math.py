"""Pins of the oracle's float packet math against libm in double precision.

The oracle restates the reference's own cephes logf (arch/NEON/
MorePacketMath.h:440-521) and psincos (MorePacketMath.h:297-393), and a
cephes expf for Eigen's pexp. Each is checked against the correctly rounded
result over the input ranges the samplers feed it; a dropped term, a wrong
constant or a swapped branch breaks these bounds by orders of magnitude.
"""
import math

import numpy as np


def ulp_err(got, exact):
    got = got.astype(np.float64)
    ulp = np.spacing(np.abs(exact).astype(np.float32)).astype(np.float64)
    return np.abs(got - exact) / ulp


def test_plog_ulp(orc):
    rng = np.random.default_rng(1)
    # the Box-Muller input 1-U is k*2^-23 in (0,1]; also wide range inputs
    k = rng.integers(1, 1 << 23, 200000)
    x = np.concatenate([(k / 2.0 ** 23).astype(np.float32),
                        np.exp(rng.uniform(-80, 80, 100000)).astype(np.float32),
                        np.float32([1.0, 2.0, 0.5, 1e-30, 3e38])])
    y = orc.plog(x)
    err = ulp_err(y, np.log(x.astype(np.float64)))
    assert err.max() <= 2.0, err.max()


def test_plog_special_values(orc):
    y = orc.plog(np.float32([0.0, -0.0, -1.0, np.inf, np.nan, 1.0]))
    assert y[0] == -np.inf and y[1] == -np.inf
    assert np.isnan(y[2]) and np.isnan(y[4])
    assert y[5] == 0.0


def test_pexp_ulp(orc):
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(-87, 88, 200000), rng.uniform(-1, 1, 100000)]).astype(np.float32)
    y = orc.pexp(x)
    err = ulp_err(y, np.exp(x.astype(np.float64)))
    assert err.max() <= 2.0, err.max()
    assert orc.pexp(np.float32([0.0]))[0] == 1.0


def test_psincos_accuracy(orc):
    rng = np.random.default_rng(3)
    # Box-Muller angles 2*pi*U, Cauchy/Poisson angles pi*U and pi*(U-1/2)
    u = (rng.integers(0, 1 << 23, 300000) / 2.0 ** 23).astype(np.float32)
    x = np.concatenate([np.float32(2 * math.pi) * u, np.float32(math.pi) * u,
                        np.float32(math.pi) * (u - np.float32(0.5))]).astype(np.float32)
    s, c = orc.psincos(x)
    xd = x.astype(np.float64)
    # cephes sinf/cosf: absolute error a few float ulps of 1
    assert np.abs(s - np.sin(xd)).max() < 3e-7
    assert np.abs(c - np.cos(xd)).max() < 3e-7
    # relative accuracy away from the zeros
    big = np.abs(np.sin(xd)) > 1e-3
    assert ulp_err(s[big], np.sin(xd[big])).max() <= 4


def test_psincos_exact_points(orc):
    s, c = orc.psincos(np.float32([0.0, -0.0]))
    assert s[0] == 0.0 and c[0] == 1.0
    assert np.signbit(s[1])  # sin(-0) keeps the sign


def test_plgamma_approx_matches_lgamma(orc):
    # lnGamma(x) ~ (x+2.5)ln(x+3) - (x+3) + ln(2pi)/2 + 1/(12(x+3)) - ln(x(x+1)(x+2))
    x = np.float32([1, 2, 3, 5, 10, 25, 100, 1000, 1e4, 1e5])
    got = orc.plgamma(x).astype(np.float64)
    exact = np.array([math.lgamma(float(v)) for v in x])
    assert np.all(np.abs(got - exact) <= 1e-4 * np.maximum(1, np.abs(exact)) + 2e-3), got - exact


def test_bit_to_ur_float_closed_form(orc):
    w = np.uint32([0, 1, 0x7FFFFF, 0x800000, 0xFFFFFFFF, 0x12345678])
    got = orc.bit_to_ur_float(w)
    exp = ((w & 0x7FFFFF).astype(np.float64) / 2 ** 23).astype(np.float32)
    np.testing.assert_array_equal(got, exp)


# ----------------------------------------------------------------- doubles ---
# The reference's double log (arch/AVX/MorePacketMath.h:781-858) and double
# sincos (MorePacketMath.h:396-491) evaluate the single-precision cephes
# polynomials in double: about 1e-9 absolute from libm, far from both a float
# result (1e-7) and a broken term (>1e-6).

def test_plog_f64_bound(orc):
    rng = np.random.default_rng(11)
    # Box-Muller / exponential input 1-U = k*2^-52 in (0,1], and a wide range
    k = rng.integers(1, 1 << 52, 300000, dtype=np.int64)
    x = np.concatenate([k / 2.0 ** 52, np.exp(rng.uniform(-700, 700, 100000)),
                        2.0 ** -np.arange(0, 53), [1.0 - 2.0 ** -52, 2.0 ** -1022]])
    y = orc.plog_f64(x)
    ref = np.log(x)
    err = np.abs(y - ref) / np.maximum(1.0, np.abs(ref))
    assert err.max() < 2e-9, err.max()
    # values within 1e-9 but not identical to libm: the float polynomial
    assert err.max() > 1e-12


def test_plog_f64_special_values(orc):
    y = orc.plog_f64(np.array([0.0, -0.0, -1.0, np.nan, 1.0, 0.5]))
    assert y[0] == -np.inf and y[1] == -np.inf
    assert np.isnan(y[2]) and np.isnan(y[3])
    assert y[4] == 0.0
    # ln 2 = q2 + q1 with q1 = -2.12194440e-4 and q2 = 0.693359375
    assert y[5] == -(0.693359375 + -2.12194440e-4)


def test_psincos_f64_bound(orc):
    rng = np.random.default_rng(12)
    u = rng.integers(0, 1 << 52, 300000, dtype=np.int64) / 2.0 ** 52
    x = np.concatenate([2 * math.pi * u, math.pi * (u - 0.5), rng.uniform(-50, 50, 100000)])
    s, c = orc.psincos_f64(x)
    assert np.abs(s - np.sin(x)).max() < 4e-9
    assert np.abs(c - np.cos(x)).max() < 4e-9
    # odd/even symmetry holds bit for bit: the sign is handled on |x|
    s2, c2 = orc.psincos_f64(-x)
    assert np.array_equal(s2, -s) and np.array_equal(c2, c)


def test_psincos_f64_exact_points(orc):
    s, c = orc.psincos_f64(np.array([0.0, -0.0]))
    assert s[0] == 0.0 and c[0] == 1.0 and c[1] == 1.0
    assert np.signbit(s[1])
