"""The batched-multivariate and serving configurations at their own sizes.

- Wishart at d = 100, 200, 256 and Dirichlet at K = 257, 1000 against the
  FP64 oracle (the largest dimensions the batched configuration draws).
- The full batched configuration in one launch per kind (4096 MvNormal with
  d ~ U{4..256} and 1024 samples, 512 Wishart with df = d + 3 and 64 draws,
  2048 Dirichlet with K ~ U{2..1000}, alpha ~ logU(0.01, 10) and 1024
  samples; parameters and engines as the benchmark draws them), with sampled
  instances, the largest among them, checked against the oracle.
- Serving requests at the configuration's extremes (multinomial K = 10^4 with
  10^6 trials, MvNormal d = 512, a 10^8-element normal request) equal to the
  direct batch calls on the client's engines.

Bounds as in test_gpu_mv.py: split-BF16 L*Z per element
(3 * 2^-16 + (d + 2) * 2^-24) * (|L||Z|)_ij; Wishart 2^-12 * (|T||T|^T)_ij;
Dirichlet (K + 1) float ulps relative.
"""
import ctypes as C

import numpy as np
import pytest

from test_gpu_fixed import assert_same
from test_gpu_mv import lz_bound, rand_cov

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def check_wishart(w_gpu, w64, a, scale, d, samples):
    t = np.abs(np.linalg.cholesky(scale)) @ np.abs(a.astype(np.float64))
    for j in range(samples):
        tj = t[:, j * d:(j + 1) * d]
        got = w_gpu[:, j * d:(j + 1) * d]
        assert np.all(np.abs(got - w64[:, j * d:(j + 1) * d]) <= 2.0 ** -12 * (tj @ tj.T)), (d, j)
        assert np.array_equal(got, got.T)


def check_dirichlet(r, r64, k):
    assert np.all(np.abs(r - r64) <= (k + 1) * 2.0 ** -23 * np.abs(r64) + 1e-38), k
    assert np.allclose(r.sum(axis=0), 1, atol=(k + 1) * 2.0 ** -22)


def check_mvnormal(x, x64, z, cov, d):
    scale = np.abs(np.linalg.cholesky(cov)) @ np.abs(z.astype(np.float64))
    err = np.abs(x - x64) / scale
    assert err.max() <= lz_bound(d), (d, err.max())


def test_wishart_largest_dims_against_fp64(er, orc):
    import orc_mv
    rng = np.random.default_rng(31)
    dims = [100, 200, 256]
    scales = [rand_cov(rng, d) for d in dims]
    df = [d + 3 for d in dims]
    samples = 8  # a multiple of the 8-float packet (the chi-square fills run over samples)
    seeds = [90, 94, 98]
    ws = er.wishart_batch(er.Vmt19937_64(seeds=seeds), scales, df, samples)
    for i, d in enumerate(dims):
        w64, a = orc_mv.wishart(orc.Engines(seeds=[seeds[i]]), df[i], scales[i], samples)
        check_wishart(ws[i].cpu().numpy().astype(np.float64), w64, a, scales[i], d, samples)


def test_dirichlet_largest_k_against_oracle(er, orc):
    import orc_mv
    rng = np.random.default_rng(32)
    ks = [257, 1000]
    alphas = [np.exp(rng.uniform(np.log(0.01), np.log(10), k)).astype(np.float32) for k in ks]
    seeds = [12, 16]
    res = er.dirichlet_batch(er.Vmt19937_64(seeds=seeds), alphas, 64)
    for i, k in enumerate(ks):
        r64, _ = orc_mv.dirichlet(orc.Engines(seeds=[seeds[i]]), alphas[i], 64)
        check_dirichlet(res[i].cpu().numpy().astype(np.float64), r64, k)


def _config4():
    """Parameters of the batched configuration, drawn as bench.py draws them."""
    rng = np.random.default_rng(4)
    dims = rng.integers(4, 257, 4096)
    covs = []
    for d in dims:
        a = rng.normal(0, 1 / np.sqrt(d), (d, d))
        covs.append(a @ a.T + d * np.eye(d))
    wd = rng.integers(4, 257, 512)
    scales = []
    for d in wd:
        a = rng.normal(0, 1 / np.sqrt(d), (d, d))
        scales.append(a @ a.T + d * np.eye(d))
    K = rng.integers(2, 1001, 2048)
    alphas = [np.exp(rng.uniform(np.log(0.01), np.log(10), k)).astype(np.float32) for k in K]
    return dims, covs, wd, scales, K, alphas


def _sampled(sizes, n=8):
    """Instance indices: the largest, the smallest and evenly spread ones."""
    idx = {int(np.argmax(sizes)), int(np.argmin(sizes)), len(sizes) - 1}
    idx.update(int(i) for i in np.linspace(0, len(sizes) - 1, n - len(idx)))
    return sorted(idx)


def test_full_batched_configuration_sampled_instances(er, orc):
    import orc_mv
    dims, covs, wd, scales, K, alphas = _config4()
    # 4096 MvNormal, 1024 samples each, instance i on engine i (seed 4i)
    gen = er.MvNormalBatch(covs)
    xs, zs = gen.generate(er.Vmt19937_64(0, n=4096, stride=4), 1024, return_z=True)
    for i in _sampled(dims):
        x64, z = orc_mv.mvnormal(orc.Engines(seeds=[4 * i]), covs[i], 1024)
        assert_same(zs[i].t().contiguous().cpu().numpy(), z.T)
        check_mvnormal(xs[i].cpu().numpy().astype(np.float64), x64, z, covs[i], int(dims[i]))
    del gen, xs, zs
    torch.cuda.empty_cache()
    # 512 Wishart, df = d + 3, 64 draws each, engines seeded 2^20 + 4i
    ws = er.WishartBatch(scales, wd + 3).generate(er.Vmt19937_64(1 << 20, n=512, stride=4), 64)
    for i in _sampled(wd, 6):
        w64, a = orc_mv.wishart(orc.Engines(seeds=[(1 << 20) + 4 * i]), int(wd[i]) + 3, scales[i], 64)
        check_wishart(ws[i].cpu().numpy().astype(np.float64), w64, a, scales[i], int(wd[i]), 64)
    del ws
    torch.cuda.empty_cache()
    # 2048 Dirichlet, 1024 samples each, engines seeded 2^21 + 4i
    res = er.dirichlet_batch(er.Vmt19937_64(1 << 21, n=2048, stride=4), alphas, 1024)
    for i in _sampled(K):
        r64, _ = orc_mv.dirichlet(orc.Engines(seeds=[(1 << 21) + 4 * i]), alphas[i], 1024)
        check_dirichlet(res[i].cpu().numpy().astype(np.float64), r64, int(K[i]))


def test_serving_at_configuration_extremes():
    import eigenrand_amd as er
    from eigenrand_amd.serving import MULTINOMIAL, MVNORMAL, spread
    from test_gpu_serve import Client, _srv
    E = 4096
    rng = np.random.default_rng(33)
    w = rng.uniform(0, 1, 10_000)
    d = 512
    cov = rand_cov(rng, d)
    n_big = 100_000_000
    with _srv(max_clients=4, engines_per_client=E, hbm_budget_bytes=8 << 30) as s:
        tm = s.submit(1, MULTINOMIAL, 16, a=1_000_000, weights=w)
        tv = s.submit(2, MVNORMAL, 2000, cov=cov)
        tn = s.submit(3, "std_normal", n_big)
        gm = s.fetch(tm, np.int32)
        gv = s.fetch(tv, np.float32)
        ptr, n_out = s.wait(tn)
        e_n, per_n = spread("std_normal", n_big, E, 8)
        assert n_out == e_n * per_n
        blocks = sorted({0, 1, e_n // 2, e_n - 1})
        host = np.empty(n_out, dtype=np.float32)
        er._lib.call("erd_memcpy_d2h", host.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), host.nbytes, None)
        for t in (tm, tv, tn):
            s.release(t)
    # multinomial K = 10^4, 10^6 trials: the conditional-binomial rows
    e, p = spread(MULTINOMIAL, 16, E, 8, a=1_000_000, dim=len(w))
    want = []
    for v in Client(1, E).parts(e):
        ms = er.multinomial_batch(v, [1_000_000] * len(v), [w] * len(v), [p] * len(v))
        want += [m.t().contiguous().cpu().numpy().ravel() for m in ms]
    np.testing.assert_array_equal(gm, np.concatenate(want))
    assert (gm.reshape(-1, len(w)).sum(axis=1) == 1_000_000).all()
    # MvNormal d = 512
    e, p = spread(MVNORMAL, 2000, E, 8, dim=d)
    want = []
    for v in Client(2, E).parts(e):
        xs = er.mvnormal_batch(v, [cov] * len(v), [p] * len(v))
        want += [x.t().contiguous().cpu().numpy().ravel() for x in xs]
    np.testing.assert_array_equal(gv, np.concatenate(want))
    # 10^8 normals over e_n engines: sampled engine blocks
    eng = Client(3, E).eng
    for g in blocks:
        ref = er.normal(per_n, 1, eng.view(g, 1)).reshape(-1).cpu().numpy()
        np.testing.assert_array_equal(host[g * per_n:(g + 1) * per_n], ref)
