"""GPU parity of the SSE2/NEON (W = 4) and AVX-512 (W = 16) lane layouts:
every fill with a form at the layout gives the oracle's outputs and engine
records bit for bit, over consecutive fills (so a fill starts mid-block and
after twists), with ragged engine counts and trailing half Box-Muller units;
calls without a form at the layout refuse the set."""
import numpy as np
import pytest

from test_gpu_fixed import assert_same, flat

pytestmark = pytest.mark.gpu

SEEDS = [5, 2 ** 33 + 1, 16380, 42, 999]

F32 = [("uniform", 0.0, 1.0, 0), ("uniform", -2.0, 3.0, 1), ("normal", 0.0, 1.0, 0), ("normal", 1.5, 2.5, 1),
       ("normal", 0.2, 0.7, 3), ("exponential", 1.7, 0.0, 0), ("cauchy", 1.0, 2.0, 0),
       ("weibull", 1.7, 2.5, 0), ("extreme_value", 0.5, 2.0, 0), ("studentt", 5.0, 0.0, 0)]


def gpu_f32(er, urng, kind, a, b, mode, per):
    if kind == "uniform":
        return er.uniform_real(per, 1, urng, a, b) if mode else er.uniform_real(per, 1, urng)
    if kind == "normal":
        if mode == 3:
            return er.lognormal(per, 1, urng, a, b)
        return er.normal(per, 1, urng, a, b) if mode else er.normal(per, 1, urng)
    if kind == "exponential":
        return er.exponential(per, 1, urng, a)
    if kind == "cauchy":
        return er.cauchy(per, 1, urng, a, b)
    if kind == "weibull":
        return er.weibull(per, 1, urng, a, b)
    if kind == "extreme_value":
        return er.extreme_value(per, 1, urng, a, b)
    return er.student_t(per, 1, urng, a)


@pytest.mark.parametrize("W", [4, 16])
@pytest.mark.parametrize("kind,a,b,mode", F32)
def test_f32_layout_fills_match_oracle(er, orc, W, kind, a, b, mode):
    urng = er.Vmt19937_64(seeds=SEEDS, layout=W)
    assert urng.layout == W
    ref = orc.Engines(seeds=SEEDS)
    # 3*W: a trailing half Box-Muller unit; 1000 packets: several twists
    for per in (3 * W, 1000 * W, W):
        got = gpu_f32(er, urng, kind, a, b, mode, per)
        assert_same(flat(got), orc.fill_f32_layout(ref, kind, per, W, a, b, mode))
        np.testing.assert_array_equal(urng.snapshot(), ref.raw)


@pytest.mark.parametrize("W", [4, 16])
@pytest.mark.parametrize("kind,a,b,mode", [("uniform", 0.0, 1.0, 0), ("uniform", -1.0, 4.0, 1),
                                           ("normal", 0.0, 1.0, 0), ("normal", 3.0, 0.5, 1),
                                           ("exponential", 2.0, 0.0, 0), ("cauchy", 0.5, 1.5, 0)])
def test_f64_layout_fills_match_oracle(er, orc, W, kind, a, b, mode):
    urng = er.Vmt19937_64(seeds=SEEDS, layout=W)
    ref = orc.Engines(seeds=SEEDS)
    PD = W // 2
    for per in (3 * PD, 700 * PD, PD):
        if kind == "uniform":
            got = er.uniform_real_f64(per, 1, urng, a, b) if mode else er.uniform_real_f64(per, 1, urng)
        elif kind == "normal":
            got = er.normal_f64(per, 1, urng, a, b) if mode else er.normal_f64(per, 1, urng)
        elif kind == "exponential":
            got = er.exponential_f64(per, 1, urng, a)
        else:
            got = er.cauchy_f64(per, 1, urng, a, b)
        assert_same(flat(got), orc.fill_f64_layout(ref, kind, per, W, a, b, mode))
        np.testing.assert_array_equal(urng.snapshot(), ref.raw)


@pytest.mark.parametrize("W", [4, 16])
def test_raw_layout_stream(er, orc, mt_golden, W):
    urng = er.Vmt19937_64(seeds=[42, 44], layout=W)
    ref = orc.Engines(seeds=[42, 44])
    got = er.randbits_u64(500, urng).cpu().numpy()
    want = orc.fill_u64_layout(ref, (2 if W == 4 else 4) * 500, W)
    np.testing.assert_array_equal(got.ravel(), want)
    if W == 4:  # the 2-lane engine seeded 42 is std::mt19937_64(42), (43) interleaved
        np.testing.assert_array_equal(got[0, 0::2], mt_golden["streams"][42][:500])
        np.testing.assert_array_equal(got[0, 1::2], mt_golden["streams"][43][:500])
    np.testing.assert_array_equal(urng.snapshot(), ref.raw)


def test_layout_views_copies_and_refusals(er):
    import torch
    urng = er.Vmt19937_64(0, n=6, layout=4)
    v = urng.view(2, 3)
    assert v.layout == 4
    assert urng.copy().layout == 4
    with pytest.raises(er.ErdError):
        er.gamma(64, 1, urng, 2.5, 1.0)
    with pytest.raises(er.ErdError):
        er.normal(64, 1, urng, precision="fast")
    with pytest.raises(er.ErdError):
        er.poisson(64, 1, urng, 3.0)
    with pytest.raises(er.ErdError):
        er.normal_vop(urng, torch.zeros(6 * 8, device="cuda"), torch.ones(6 * 8, device="cuda"))
    with pytest.raises(er.ErdError):  # a W = 16 fill takes whole Packet16f
        er.normal(8, 1, er.Vmt19937_64(0, n=2, layout=16))
    with pytest.raises(er.ErdError):
        urng.set_layout(5)
    urng.set_layout(8)
    er.gamma(64, 1, urng, 2.5, 1.0)
