"""The oracle's SSE2/NEON (W = 4) and AVX-512 (W = 16) lane layouts.

Pins, each independent of the layout code under test:
- at W = 8 the layout functions equal the AVX2 fills, which are pinned
  elsewhere (std streams, closed forms, stream recomputations);
- at W = 16 every element-wise fill equals the AVX2 fill bit for bit (a
  Packet16f is two consecutive calls, low half first: the same words in the
  same order, arch/AVX512/RandUtils.h:37-46);
- the W = 4 engine seeded x0 is the 2-lane Pmt19937_64<Packet4i>
  (PacketRandomEngine.h:587-588): its call k is [mt19937_64(x0)[k],
  mt19937_64(x0+1)[k]], checked against the libstdc++ golden streams;
- the Box-Muller and Student-t pairings of each layout (u1 = the first
  packet, u2 = the second; the sine packet returned and the cosine packet
  cached), recomputed in float64 from the raw words. Swapping u1 and u2, a
  wrong call order inside a Packet16f or lanes 2-3 at W = 4 all fail these.
"""
import numpy as np
import pytest

SEEDS = [42, 7, 16380]


def words32(raw_u64):
    """The 32-bit words of a u64 stream, low word first."""
    return raw_u64.view(np.uint32)


def ur(w):
    """bit_to_ur_float (MorePacketMath.h:205-215) in float64: exact."""
    return ((w & 0x7FFFFF).astype(np.float64)) / 2.0 ** 23


def packets(e_raw_calls, W):
    """Per-engine float-word packets of layout W from the AVX2 calls
    (n_calls, 4) of one engine: W = 4 takes lanes 0-1 of each call, W = 16
    joins two calls."""
    if W == 4:
        words = words32(np.ascontiguousarray(e_raw_calls[:, :2])).reshape(-1, 4)
    elif W == 8:
        words = words32(np.ascontiguousarray(e_raw_calls)).reshape(-1, 8)
    else:
        words = words32(np.ascontiguousarray(e_raw_calls)).reshape(-1, 16)
    return words


@pytest.mark.parametrize("kind,a,b,mode", [("uniform", 0.0, 1.0, 0), ("uniform", -2.0, 3.0, 1),
                                           ("normal", 0.0, 1.0, 0), ("normal", 1.5, 2.5, 1),
                                           ("normal", 0.2, 0.7, 3), ("exponential", 1.7, 0.0, 0),
                                           ("cauchy", 1.0, 2.0, 0), ("weibull", 1.7, 2.5, 0),
                                           ("extreme_value", 0.5, 2.0, 0), ("studentt", 5.0, 0.0, 0)])
def test_w8_equals_the_avx2_fills(orc, kind, a, b, mode):
    per = 4096
    e1, e2 = orc.Engines(seeds=SEEDS), orc.Engines(seeds=SEEDS)
    got = orc.fill_f32_layout(e1, kind, per, 8, a, b, mode)
    if kind == "uniform":
        want = orc.fill_uniform_f32(e2, per, a, b)
    elif kind == "normal" and mode == 3:
        want = orc.fill_lognormal_f32(e2, per, a, b)
    elif kind == "normal":
        want = orc.fill_normal_f32(e2, per, a, b) if mode else orc.fill_normal_f32(e2, per)
    elif kind == "exponential":
        want = orc.fill_exponential_f32(e2, per, a)
    elif kind == "cauchy":
        want = orc.fill_cauchy_f32(e2, per, a, b)
    elif kind == "weibull":
        want = orc.fill_weibull_f32(e2, per, a, b)
    elif kind == "extreme_value":
        want = orc.fill_extreme_value_f32(e2, per, a, b)
    else:
        want = orc.fill_studentt_f32(e2, per, a)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    np.testing.assert_array_equal(e1.raw, e2.raw)


@pytest.mark.parametrize("kind,a,b,mode", [("uniform", 0.0, 1.0, 0), ("uniform", -1.0, 4.0, 1),
                                           ("normal", 0.0, 1.0, 0), ("normal", 3.0, 0.5, 1),
                                           ("exponential", 2.0, 0.0, 0), ("cauchy", 0.5, 1.5, 0)])
def test_w8_f64_equals_the_avx2_fills(orc, kind, a, b, mode):
    per = 2048
    e1, e2 = orc.Engines(seeds=SEEDS), orc.Engines(seeds=SEEDS)
    got = orc.fill_f64_layout(e1, kind, per, 8, a, b, mode)
    if kind == "uniform":
        want = orc.fill_uniform_f64(e2, per, a, b)
    elif kind == "normal":
        want = orc.fill_normal_f64(e2, per, a, b) if mode else orc.fill_normal_f64(e2, per)
    elif kind == "exponential":
        want = orc.fill_exponential_f64(e2, per, a)
    else:
        want = orc.fill_cauchy_f64(e2, per, a, b)
    np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))
    np.testing.assert_array_equal(e1.raw, e2.raw)


@pytest.mark.parametrize("kind", ["uniform", "exponential", "cauchy", "weibull", "extreme_value"])
def test_w16_elementwise_equals_w8(orc, kind):
    per = 4096
    e1, e2 = orc.Engines(seeds=SEEDS), orc.Engines(seeds=SEEDS)
    a, b = (1.3, 2.0) if kind != "uniform" else (0.0, 1.0)
    np.testing.assert_array_equal(orc.fill_f32_layout(e1, kind, per, 16, a, b).view(np.uint32),
                                  orc.fill_f32_layout(e2, kind, per, 8, a, b).view(np.uint32))
    np.testing.assert_array_equal(e1.raw, e2.raw)
    for k64 in ("uniform", "exponential", "cauchy"):
        np.testing.assert_array_equal(orc.fill_f64_layout(e1, k64, 512, 16, 1.0, 2.0).view(np.uint64),
                                      orc.fill_f64_layout(e2, k64, 512, 8, 1.0, 2.0).view(np.uint64))
    np.testing.assert_array_equal(orc.fill_u64_layout(e1, 400, 16), orc.fill_u64_layout(e2, 400, 8))


def test_w4_engine_is_the_two_lane_std_engine(orc, mt_golden):
    # Pmt19937_64<Packet4i>(x0): lanes x0, x0+1 (PacketRandomEngine.h:230-239)
    for x0 in (42, 44, 0, 16382):
        e = orc.Engines(seeds=[x0])
        d = orc.fill_u64_layout(e, 2000, 4).reshape(1000, 2)
        np.testing.assert_array_equal(d[:, 0], mt_golden["streams"][x0])
        np.testing.assert_array_equal(d[:, 1], mt_golden["streams"][x0 + 1])


def test_w4_uniform_is_the_words_of_each_call(orc, mt_golden):
    e = orc.Engines(seeds=[42])
    u = orc.fill_f32_layout(e, "uniform", 4000, 4)
    s0 = mt_golden["streams"][42].astype(np.uint64)
    s1 = mt_golden["streams"][43].astype(np.uint64)
    calls = np.stack([s0, s1], axis=1)                    # (1000, 2) u64
    w = words32(np.ascontiguousarray(calls)).reshape(-1)  # float j of call k = word j
    np.testing.assert_array_equal(u.astype(np.float64), ur(w))


def raw_calls(orc, seeds, n_calls):
    return orc.fill_u64(orc.Engines(seeds=seeds), n_calls).reshape(len(seeds), n_calls, 4)


@pytest.mark.parametrize("W", [4, 16])
def test_box_muller_pairing(orc, W):
    per = 2048
    calls = raw_calls(orc, SEEDS, 4 * per)
    z = orc.fill_f32_layout(orc.Engines(seeds=SEEDS), "normal", per, W).reshape(len(SEEDS), per)
    for g in range(len(SEEDS)):
        pk = packets(calls[g], W)                         # packets of W words
        out = np.empty(per)
        for m in range(per // (2 * W)):
            u1, u2 = 1.0 - ur(pk[2 * m]), ur(pk[2 * m + 1])
            r = np.sqrt(-2.0 * np.log(u1))
            out[2 * W * m:2 * W * m + W] = r * np.sin(2 * np.pi * u2)
            out[2 * W * m + W:2 * W * (m + 1)] = r * np.cos(2 * np.pi * u2)
        np.testing.assert_allclose(z[g], out, rtol=2e-6, atol=2e-6)


@pytest.mark.parametrize("W", [4, 16])
def test_student_t_pairing(orc, W):
    per, nu = 2048, 5.0
    calls = raw_calls(orc, SEEDS, 4 * per)
    t = orc.fill_f32_layout(orc.Engines(seeds=SEEDS), "studentt", per, W, nu).reshape(len(SEEDS), per)
    for g in range(len(SEEDS)):
        pk = packets(calls[g], W)
        out = np.empty(per)
        for m in range(per // W):
            u1, u2 = 1.0 - ur(pk[2 * m]), ur(pk[2 * m + 1])
            r = np.sqrt(nu * (np.exp(np.log(u1) * (-2 / nu)) - 1.0))
            out[W * m:W * (m + 1)] = r * np.sin(2 * np.pi * u2)
        np.testing.assert_allclose(t[g], out, rtol=2e-5, atol=2e-5)


@pytest.mark.parametrize("W", [4, 16])
def test_normal_f64_pairing(orc, W):
    per = 1024
    PD = W // 2
    calls = raw_calls(orc, SEEDS, 4 * per)
    z = orc.fill_f64_layout(orc.Engines(seeds=SEEDS), "normal", per, W).reshape(len(SEEDS), per)
    for g in range(len(SEEDS)):
        c = calls[g][:, :2] if W == 4 else calls[g]
        u = ((c & np.uint64(2 ** 52 - 1)).astype(np.float64) / 2.0 ** 52).reshape(-1, PD)  # bit_to_ur_double
        out = np.empty(per)
        for m in range(per // (2 * PD)):
            u1, u2 = 1.0 - u[2 * m], u[2 * m + 1]
            r = np.sqrt(-2.0 * np.log(u1))
            out[2 * PD * m:2 * PD * m + PD] = r * np.sin(2 * np.pi * u2)
            out[2 * PD * m + PD:2 * PD * (m + 1)] = r * np.cos(2 * np.pi * u2)
        np.testing.assert_allclose(z[g], out, rtol=1e-8, atol=1e-8)


@pytest.mark.parametrize("kind", ["exponential", "cauchy", "weibull", "extreme_value"])
def test_w4_elementwise_fills_follow_the_w4_words(orc, kind):
    """Each output mapped back through its distribution's inverse is the W = 4
    uniform of the same position (the inverse maps are well conditioned where
    the forward ones amplify float rounding, e.g. the Cauchy tails)."""
    per = 4096
    a, b = 1.5, 2.0
    v = orc.fill_f32_layout(orc.Engines(seeds=SEEDS), kind, per, 4, a, b).astype(np.float64)
    u = orc.fill_f32_layout(orc.Engines(seeds=SEEDS), "uniform", per, 4).astype(np.float64)
    if kind == "exponential":
        back = 1 - np.exp(-a * v)
    elif kind == "cauchy":
        back = np.arctan((v - a) / b) / np.pi + 0.5
    elif kind == "weibull":
        back = 1 - np.exp(-(v / b) ** a)
    else:
        back = np.exp(-np.exp((a - v) / b)) - 2.0 ** -26
    np.testing.assert_allclose(back, u, rtol=0, atol=3e-6)


def test_w4_f64_is_lanes_0_1(orc):
    per = 1000
    calls = raw_calls(orc, SEEDS, per // 2)
    got = orc.fill_f64_layout(orc.Engines(seeds=SEEDS), "uniform", per, 4).reshape(len(SEEDS), per)
    want = (calls[:, :, :2] & np.uint64(2 ** 52 - 1)).astype(np.float64).reshape(len(SEEDS), per) / 2.0 ** 52
    np.testing.assert_array_equal(got, want)
