"""The oracle's engine model against the C++ standard engine.

Every lane of a Vmt19937_64 is std::mt19937_64(seed + lane)
(PacketRandomEngine.h:230-265, 295-310); the golden streams come from
libstdc++ (tests/golden/gen_golden.py).
"""
import numpy as np


def test_known_answer_default_seed(orc, mt_golden):
    # [rand.predef]: 10000th draw of a default-constructed mt19937_64.
    assert mt_golden["kat"] == 9981545732273789042
    e = orc.Engines(seeds=[5489])
    draws = orc.fill_u64(e, 10000).reshape(10000, 4)
    assert int(draws[9999, 0]) == 9981545732273789042
    np.testing.assert_array_equal(draws[:1000, 0], mt_golden["streams"][5489])


def test_lanes_are_std_streams(orc, mt_golden):
    e = orc.Engines(seeds=[42])
    d = orc.fill_u64(e, 1000).reshape(1000, 4)
    for lane in range(4):
        np.testing.assert_array_equal(d[:, lane], mt_golden["streams"][42 + lane])


def test_range_engines_and_block_split(orc, mt_golden):
    # engines seeded 0 and 4 (= lanes 0..7), drawn in two pieces that cross the
    # 312-word block boundary at odd offsets
    e = orc.Engines(n=2, seed0=0, stride=4)
    a = orc.fill_u64(e, 311).reshape(2, 311, 4)
    b = orc.fill_u64(e, 689).reshape(2, 689, 4)
    d = np.concatenate([a, b], axis=1)
    for g in range(2):
        for lane in range(4):
            np.testing.assert_array_equal(d[g, :, lane], mt_golden["streams"][4 * g + lane])


def test_last_engines_of_config2(orc, mt_golden):
    e = orc.Engines(seeds=[16380])
    d = orc.fill_u64(e, 1000).reshape(1000, 4)
    for lane in range(4):
        np.testing.assert_array_equal(d[:, lane], mt_golden["streams"][16380 + lane])


def test_half_serves_low_then_cached_high(orc, mt_golden):
    # half() (PacketRandomEngine.h:325-336): low 128 bits, then cached high.
    e = orc.Engines(seeds=[42])
    h = orc.fill_half_u64(e, 8).reshape(8, 2)
    s = mt_golden["streams"]
    for k in range(4):
        np.testing.assert_array_equal(h[2 * k], [s[42][k], s[43][k]])
        np.testing.assert_array_equal(h[2 * k + 1], [s[44][k], s[45][k]])


def test_p8_adaptor_lane_order(orc, mt_golden):
    # P8_mt19937_64(seed): out[8k + j] = mt19937_64(seed + j)[k]
    # (ParallelRandomEngineAdaptor, PacketRandomEngine.h:437-444, 487-494).
    e = orc.Engines(n=2, seed0=0, stride=4)
    out = orc.fill_p8_u64(e, 50).reshape(50, 8)
    for j in range(8):
        np.testing.assert_array_equal(out[:, j], mt_golden["streams"][j][:50])


def test_uniform_is_closed_form_of_stream(orc, mt_golden):
    # bit_to_ur_float (MorePacketMath.h:205-215): float j of a call is word j.
    e = orc.Engines(seeds=[42])
    u = orc.fill_uniform_f32(e, 8 * 100).reshape(100, 8)
    s = mt_golden["streams"]
    for j in range(8):
        w = s[42 + j // 2][:100]
        word = (w >> np.uint64(32 * (j % 2))) & np.uint64(0xFFFFFFFF)
        exp = ((word & np.uint64(0x7FFFFF)).astype(np.float64) / 2.0 ** 23).astype(np.float32)
        np.testing.assert_array_equal(u[:, j], exp)


def test_uniform_f64_closed_form(orc, mt_golden):
    e = orc.Engines(seeds=[0])
    u = orc.fill_uniform_f64(e, 4 * 64).reshape(64, 4)
    for lane in range(4):
        w = mt_golden["streams"][lane][:64]
        exp = (w & np.uint64((1 << 52) - 1)).astype(np.float64) / 2.0 ** 52
        np.testing.assert_array_equal(u[:, lane], exp)


def test_discard_equals_drawing(orc):
    e1, e2 = orc.Engines(seeds=[9, 10]), orc.Engines(seeds=[9, 10])
    orc.discard(e1, 1234)
    orc.fill_u64(e2, 1234)
    np.testing.assert_array_equal(e1.raw, e2.raw)
