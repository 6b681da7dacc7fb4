"""The fast precision mode of normal<float> and the store-ceiling kernel.

The fast mode consumes exactly the draws of the bit-exact mode (so engine
records must be identical after either), pairs them the same way and writes
the same output order; only log, sqrt, sin and cos come from the hardware.
Stated bound per output z, against the bit-exact value z_e of the same slot:

    |z - z_e| <= 4 ulp(z_e) + 2^-20 * r,   r = sqrt(z_e^2 + z_e'^2) (<= 5.65)

where z_e' is the other half (sin/cos) of the slot's Box-Muller pair. The
bit-exact mode is pinned to the oracle elsewhere (test_gpu_fixed.py), and is
checked against it here again on the small cases.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOUND_ULP = 4.0
BOUND_ABS_PER_R = 2.0 ** -20


def pair_radius(z):
    """r of each slot's Box-Muller pair: blocks of 16 = [r sin x8, r cos x8]."""
    z = np.asarray(z, np.float64).reshape(-1, 2, 8)
    r = np.sqrt(z[:, 0, :] ** 2 + z[:, 1, :] ** 2)
    return np.repeat(r[:, None, :], 2, axis=1).reshape(-1)


def ulp32(x):
    a = np.abs(np.asarray(x, np.float32))
    return (np.nextafter(a, np.float32(np.inf)) - a).astype(np.float64)


def check_bound(fast, exact, scale=1.0, affine_exact=None):
    """fast, exact: standard normals (or, with affine_exact, z*sd+mean where
    `exact` holds the standard values that give the radius)."""
    fast = np.asarray(fast, np.float64)
    r = pair_radius(exact)
    if affine_exact is None:
        ref = np.asarray(exact, np.float64)
        lim = BOUND_ULP * ulp32(exact) + BOUND_ABS_PER_R * r
    else:
        ref = np.asarray(affine_exact, np.float64)
        lim = scale * (BOUND_ULP * ulp32(exact) + BOUND_ABS_PER_R * r) + 2 * ulp32(affine_exact)
    err = np.abs(fast - ref)
    bad = np.nonzero(err > lim)[0]
    assert bad.size == 0, f"{bad.size} outside the bound, first {bad[:4]}: {fast[bad[:4]]} vs {ref[bad[:4]]}"
    return float((err / np.maximum(lim, 1e-30)).max())


@pytest.mark.parametrize("per", [8, 16, 24, 624, 4992, 65536 + 8])
def test_fast_normal_same_draws_and_bound_vs_oracle(er, orc, per):
    seeds = [42, 7, 1000]
    fast_eng = er.Vmt19937_64(seeds=seeds)
    ex_eng = orc.Engines(seeds=seeds)
    # start both at an odd draw index so the first block boundary splits a
    # Box-Muller unit (the carried-word path)
    er.randbits_u64(5, fast_eng)
    orc.fill_u64(ex_eng, 5)
    fast = er.normal(per, 1, fast_eng, precision="fast").reshape(len(seeds), per).cpu().numpy()
    exact = orc.fill_normal_f32(ex_eng, per).reshape(len(seeds), per)
    np.testing.assert_array_equal(fast_eng.snapshot(), ex_eng.raw)
    full = per - per % 16
    for g in range(len(seeds)):
        if full:
            check_bound(fast[g, :full], exact[g, :full])
        if per % 16:
            # a trailing sine packet: its radius from the exact cos packet is
            # not written; bound it with the largest radius (5.65)
            t = slice(full, per)
            lim = BOUND_ULP * ulp32(exact[g, t]) + BOUND_ABS_PER_R * 5.66
            assert (np.abs(fast[g, t].astype(np.float64) - exact[g, t]) <= lim).all()


def test_fast_normal_affine(er, orc):
    seeds = [3, 4]
    per = 4096
    mean, sd = -2.5, 3.25
    fe = er.Vmt19937_64(seeds=seeds)
    oe = orc.Engines(seeds=seeds)
    fast = er.normal(per, 1, fe, mean, sd, precision="fast").reshape(2, per).cpu().numpy()
    z = orc.fill_normal_f32(oe, per).reshape(2, per)
    xe = (z * np.float32(sd) + np.float32(mean)).astype(np.float32)
    for g in range(2):
        check_bound(fast[g], z[g], scale=sd, affine_exact=xe[g])
    np.testing.assert_array_equal(fe.snapshot(), oe.raw)


def test_fast_normal_full_size_bound_and_moments(er, orc):
    """2^28 fast normals from 4096 engines seeded 0,4,...,16380: every output
    within the bound of the bit-exact fill of the same engines (compared on
    the GPU), the engine records identical, and the distribution checks."""
    import torch
    from scipy import stats
    n_eng, per = 4096, 65536
    a = er.Vmt19937_64(0, n=n_eng, stride=4)
    b = er.Vmt19937_64(0, n=n_eng, stride=4)
    fast = er.normal(per, 1, a, precision="fast").reshape(-1)
    exact = er.normal(per, 1, b).reshape(-1)
    assert np.array_equal(a.snapshot(), b.snapshot())
    # bound on the device, in float64
    e3 = exact.view(-1, 2, 8).double()
    r = torch.sqrt(e3[:, 0, :] ** 2 + e3[:, 1, :] ** 2).unsqueeze(1).expand(-1, 2, -1).reshape(-1)
    ae = exact.abs()
    ulp = (torch.nextafter(ae, torch.full_like(ae, float("inf"))) - ae).double()
    err = (fast.double() - exact.double()).abs()
    lim = BOUND_ULP * ulp + BOUND_ABS_PER_R * r
    worst = (err / lim).max().item()
    assert worst <= 1.0, worst
    del e3, r, ae, ulp, err, lim
    # the exact fill is pinned to the oracle: spot-check one block here too
    ref = orc.Engines(seeds=[4 * 777])
    np.testing.assert_array_equal(exact.view(n_eng, per)[777].cpu().numpy().view(np.uint32),
                                  orc.fill_normal_f32(ref, per).view(np.uint32))
    x = fast.double()
    n = x.numel()
    m = x.mean().item()
    v = x.var().item()
    k = ((x - m) ** 4).mean().item() / v ** 2
    assert abs(m) < 5 / n ** 0.5
    assert abs(v - 1) < 5 * (2 / n) ** 0.5
    assert abs(k - 3) < 5 * (24 / n) ** 0.5
    assert torch.isfinite(fast).all()
    sample = fast[:: 97].cpu().numpy().astype(np.float64)
    assert stats.kstest(sample, "norm").pvalue > 1e-4


def test_store_ceiling_writes_exactly_its_blocks(er):
    import torch
    n_blocks, per = 37, 1024
    out = torch.zeros(n_blocks * per + 64, dtype=torch.float32, device="cuda")
    er.store_ceiling(n_blocks, per, out)
    torch.cuda.synchronize()
    assert bool((out[: n_blocks * per] == 1.0).all())
    assert bool((out[n_blocks * per:] == 0.0).all())
    with pytest.raises(er.InvalidArgument):
        er.store_ceiling(2, 6, out)  # not a multiple of 4 floats
