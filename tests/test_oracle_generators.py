"""Reused generator objects in the oracle (Basic.h:268-342): consecutive
generate() calls on one generator continue exactly where a single larger call
would be, i.e. the cached normal cos packet and the CompressMask leftovers
carry over, while fresh generators (the free functions) drop them."""
import numpy as np
import pytest

CASES = [
    (3, 0.0, 0.0, 8, [8, 24, 40, 8]),          # std normal: odd packet counts leave a cos cached
    (4, 2.0, 0.5, 8, [24, 8, 16]),             # normal(mean, sd)
    (16, 0.2, 0.7, 8, [8, 8, 24]),             # lognormal: NormalGen's cache
    (9, 0.4, 1.5, 8, [8, 16, 40, 8]),          # gamma GS (alpha < 1)
    (9, 3.7, 1.0, 8, [16, 8, 24]),             # gamma tangent
    (9, 3.0, 2.0, 8, [8, 16]),                 # integer alpha: no state
    (10, 30.0, 0.0, 4, [4, 12, 20, 4]),        # Poisson tangent
    (10, 2.5, 0.0, 4, [4, 8, 12]),             # Poisson Knuth: no state
    (11, 100, 0.3, 4, [4, 12, 8]),             # binomial tangent
    (11, 10, 0.6, 4, [8, 4]),                  # binomial n < 25: no state
]


def plain(orc, e, dist, a, b, n):
    if dist == 3:
        return orc.fill_normal_f32(e, n)
    if dist == 4:
        return orc.fill_normal_f32(e, n, a, b)
    if dist == 16:
        return orc.fill_lognormal_f32(e, n, a, b)
    if dist == 9:
        return orc.fill_gamma_f32(e, n, a, b)
    if dist == 10:
        return orc.fill_poisson_i32(e, n, a)
    return orc.fill_binomial_i32(e, n, int(a), b)


@pytest.mark.parametrize("dist,a,b,w,sizes", CASES)
def test_generator_calls_concatenate(orc, dist, a, b, w, sizes):
    seeds = [3, 77, 2 ** 40 + 5]
    e1 = orc.Engines(seeds=seeds)
    gen = orc.Generators(dist, a, b, len(seeds))
    parts = [gen.fill(e1, n).reshape(len(seeds), n) for n in sizes]
    e2 = orc.Engines(seeds=seeds)
    whole = plain(orc, e2, dist, a, b, sum(sizes)).reshape(len(seeds), -1)
    np.testing.assert_array_equal(np.concatenate(parts, axis=1), whole)
    np.testing.assert_array_equal(e1.raw, e2.raw)


def test_fresh_generators_drop_leftovers(orc):
    """The free functions start from an empty cache each call, so split
    calls differ from the reused generator once a leftover exists."""
    seeds = [11]
    gen = orc.Generators(3, 0, 0, 1)
    e1, e2 = orc.Engines(seeds=seeds), orc.Engines(seeds=seeds)
    g = np.concatenate([gen.fill(e1, 8), gen.fill(e1, 8)])
    f = np.concatenate([orc.fill_normal_f32(e2, 8), orc.fill_normal_f32(e2, 8)])
    np.testing.assert_array_equal(g[:8], f[:8])
    assert not np.array_equal(g[8:], f[8:])  # the cached cos packet vs a new sin packet
