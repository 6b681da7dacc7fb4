"""Batched multivariate generators on the GPU against the FP64 oracle.

The univariate draws inside each generator are bit-exact to the oracle (the
standard normals Z of MvNormal are compared directly); the float results of
the products are bounded against float64 products of the same draws, with
the bound written in each test; distribution checks follow test_mv.cpp
(EXPECT_SIMILAR_MATRIX, 10% relative)."""
import numpy as np
import pytest

from test_gpu_fixed import assert_same

pytestmark = pytest.mark.gpu


def rand_cov(rng, d):
    a = rng.normal(0, 1 / np.sqrt(d), (d, d))
    return a @ a.T + d * np.eye(d)  # cov = A A^T + d I  (config 4)


def test_cholesky_fp64_matches_lapack(er):
    rng = np.random.default_rng(1)
    dims = [1, 4, 17, 64, 100, 256, 300]
    mats = [rand_cov(rng, d) for d in dims]
    ls, info, rc = er.cholesky_batch(mats)
    assert rc == 0 and not info.any()
    for m, l in zip(mats, ls):
        ref = np.linalg.cholesky(m)
        assert np.allclose(l, ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())
        assert np.allclose(np.triu(l, 1), 0)


def test_cholesky_reports_not_psd(er):
    m = np.diag([1.0, -1.0, 2.0])
    _, info, rc = er.cholesky_batch([np.eye(2), m])
    assert rc == 3 and info[0] == 0 and info[1] == 2


def test_cholesky_unblocked_path_above_512(er):
    # a batch whose largest d exceeds 512 takes the unblocked kernel for all
    rng = np.random.default_rng(5)
    mats = [rand_cov(rng, d) for d in (600, 33)]
    ls, info, rc = er.cholesky_batch(mats)
    assert rc == 0 and not info.any()
    for m, l in zip(mats, ls):
        ref = np.linalg.cholesky(m)
        assert np.allclose(l, ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())
        assert np.allclose(np.triu(l, 1), 0)


def test_cholesky_reports_late_pivot(er):
    # M = U D U^T with U unit lower triangular: the j-th Cholesky pivot is D[j],
    # so D[49] = -1 fails at column 50 (1-based), in the second 32-wide panel
    rng = np.random.default_rng(6)
    d = 70
    u = np.tril(rng.normal(0, 0.3, (d, d)), -1) + np.eye(d)
    dg = np.full(d, 2.0)
    dg[49] = -1.0
    m = u @ np.diag(dg) @ u.T
    _, info, rc = er.cholesky_batch([rand_cov(rng, 40), m])
    assert rc == 3 and info[0] == 0 and info[1] == 50


# split-BF16 (hi+lo) with FP32 accumulation: bf16 keeps 8 significant bits,
# so |x - hi| <= 2^-8 |x| and the hi+lo residual is <= 2^-16 |x|; the two
# residuals and the dropped lo*lo term give <= 3 * 2^-16 per product, so per
# element
#   |X - X64| <= (3 * 2^-16 + (d + 2) * 2^-24) * sum_k |L_ik||Z_kj| + 2^-22 |mean_i|
# (the (d + 2) * 2^-24 term is the FP32 accumulation). Tested bound.
def lz_bound(d):
    return 3 * 2.0 ** -16 + (d + 2) * 2.0 ** -24


@pytest.mark.parametrize("dims", [[4, 9, 16, 17, 31, 64], [65, 100, 128, 200, 256], [300, 512]])
def test_mvnormal_batch_against_fp64(er, orc, dims):
    import orc_mv
    rng = np.random.default_rng(sum(dims))
    samples = [64 if d % 2 else 40 for d in dims]
    covs = [rand_cov(rng, d) for d in dims]
    means = [rng.uniform(-10, 10, d).astype(np.float32) for d in dims]
    seeds = [1000 + 4 * i for i in range(len(dims))]
    urng = er.Vmt19937_64(seeds=seeds)
    xs, zs = er.mvnormal_batch(urng, covs, samples, means, return_z=True)
    worst = 0.0
    for i, d in enumerate(dims):
        x64, z = orc_mv.mvnormal(orc.Engines(seeds=[seeds[i]]), covs[i], samples[i], means[i])
        assert_same(zs[i].t().contiguous().cpu().numpy(), z.T)  # the normals are bit-exact
        x = xs[i].cpu().numpy().astype(np.float64)
        scale = np.abs(np.linalg.cholesky(covs[i])) @ np.abs(z.astype(np.float64))
        err = np.abs(x - x64) / (scale + np.abs(means[i])[:, None] * 2.0 ** -6 / d)
        worst = max(worst, float(err.max()))
        assert err.max() <= lz_bound(d), (d, err.max())
    print("max |X-X64| / (|L||Z|) =", worst)


def test_lz_variants_give_identical_results(er, monkeypatch):
    """The L*Z tuner may pick any of the four tile variants per dimension
    bucket: each must give the same bits (same per-element k order)."""
    rng = np.random.default_rng(21)
    dims = [3, 16, 40, 100, 130, 256, 300]
    covs = [rand_cov(rng, d) for d in dims]
    means = [rng.normal(size=d).astype(np.float32) for d in dims]
    outs = []
    for v in range(4):
        monkeypatch.setenv("ERD_LZ_VARIANT", str(v))
        xs = er.mvnormal_batch(er.Vmt19937_64(seeds=list(range(7))), covs, [40, 8, 72, 64, 24, 16, 8], means)
        outs.append([x.cpu().numpy() for x in xs])
    for v in range(1, 4):
        for a, b in zip(outs[0], outs[v]):
            assert_same(a, b)


def test_mvnormal_distribution_9d(er):
    """test_mv.cpp:98-122: 9-d normal, sample mean/cov within 10%."""
    rng = np.random.default_rng(9)
    d = 9
    mean = np.arange(d, dtype=np.float32)
    cov = rand_cov(rng, d)
    gen = er.MvNormalGen(mean, cov)
    urng = er.Vmt19937_64(42)
    x = gen.generate(urng, 8000).cpu().numpy().astype(np.float64)
    assert np.allclose(x.mean(axis=1), mean, atol=0.1 * np.sqrt(np.diag(cov)).max())
    assert np.allclose(np.cov(x), cov, rtol=0.1, atol=0.1 * np.abs(cov).max())


def test_wishart_batch_against_fp64(er, orc):
    import orc_mv
    rng = np.random.default_rng(3)
    dims = [4, 7, 33, 64]
    scales = [rand_cov(rng, d) for d in dims]
    df = [d + 3 for d in dims]
    samples = [16, 8, 8, 8]
    seeds = [50, 60, 70, 80]
    urng = er.Vmt19937_64(seeds=seeds)
    ws = er.wishart_batch(urng, scales, df, samples)
    for i, d in enumerate(dims):
        w64, a = orc_mv.wishart(orc.Engines(seeds=[seeds[i]]), df[i], scales[i], samples[i])
        w = ws[i].cpu().numpy().astype(np.float64)
        # bound: |W - W64| <= 2^-12 * (|T| |T|^T) with |T| = |L| |A|
        t = np.abs(np.linalg.cholesky(scales[i])) @ np.abs(a.astype(np.float64))
        for j in range(samples[i]):
            tj = t[:, j * d:(j + 1) * d]
            bound = 2.0 ** -12 * (tj @ tj.T)
            assert np.all(np.abs(w[:, j * d:(j + 1) * d] - w64[:, j * d:(j + 1) * d]) <= bound), (d, j)
            assert np.array_equal(w[:, j * d:(j + 1) * d], w[:, j * d:(j + 1) * d].T)  # symmetric


def test_invwishart_batch_against_fp64(er, orc):
    """InvWishartGen: W = U^T U with U = (L A)^-1, L = chol(scale^-1). The
    FP32 triangular inverse amplifies the split-BF16 error of T by the
    condition of T; with these scales the relative error stays below 2^-10."""
    import orc_mv
    rng = np.random.default_rng(8)
    dims = [4, 7, 33, 64, 100]
    scales = [rand_cov(rng, d) for d in dims]
    df = [d + 10 for d in dims]
    samples = [16, 8, 8, 8, 8]
    seeds = [51, 61, 71, 81, 91]
    urng = er.Vmt19937_64(seeds=seeds)
    ws = er.invwishart_batch(urng, scales, df, samples)
    for i, d in enumerate(dims):
        w64, _ = orc_mv.invwishart(orc.Engines(seeds=[seeds[i]]), df[i], scales[i], samples[i])
        w = ws[i].cpu().numpy().astype(np.float64)
        for j in range(samples[i]):
            ref = w64[:, j * d:(j + 1) * d]
            got = w[:, j * d:(j + 1) * d]
            assert np.abs(got - ref).max() <= 2.0 ** -10 * np.abs(ref).max(), (d, j)
            assert np.array_equal(got, got.T)


def test_invwishart_large_dims(er, orc):
    """Panels up to 16 blocks (d = 300, 512) next to small ones in one launch."""
    import orc_mv
    rng = np.random.default_rng(12)
    dims = [300, 5, 512]
    scales = [rand_cov(rng, d) for d in dims]
    df = [d + 10 for d in dims]
    seeds = [31, 41, 51]
    ws = er.invwishart_batch(er.Vmt19937_64(seeds=seeds), scales, df, 8)
    for i, d in enumerate(dims):
        w64, _ = orc_mv.invwishart(orc.Engines(seeds=[seeds[i]]), df[i], scales[i], 8)
        w = ws[i].cpu().numpy().astype(np.float64)
        for j in (0, 7):
            ref = w64[:, j * d:(j + 1) * d]
            got = w[:, j * d:(j + 1) * d]
            assert np.abs(got - ref).max() <= 2.0 ** -10 * np.abs(ref).max(), (d, j)


def test_invwishart_mean(er):
    """E[W] = scale / (df - d - 1) for the inverse Wishart, within 10%."""
    rng = np.random.default_rng(9)
    d, df, n = 4, 12, 8192
    scale = rand_cov(rng, d)
    w = er.InvWishartGen(df, scale).generate(er.Vmt19937_64(13), n).cpu().numpy().astype(np.float64)
    mean = w.reshape(d, n, d).mean(axis=1)
    assert np.allclose(mean, scale / (df - d - 1), rtol=0.1, atol=0.1 * np.abs(scale).max() / (df - d - 1))


def test_wishart_mean(er):
    """test_mv.cpp:124-150: E[W] = df * scale within 10%."""
    rng = np.random.default_rng(4)
    d, df, n = 5, 8, 4096
    scale = rand_cov(rng, d)
    w = er.WishartGen(df, scale).generate(er.Vmt19937_64(7), n).cpu().numpy().astype(np.float64)
    mean = w.reshape(d, n, d).mean(axis=1)
    assert np.allclose(mean, df * scale, rtol=0.1, atol=0.1 * df * np.abs(scale).max())


def test_dirichlet_batch_against_oracle(er, orc):
    import orc_mv
    rng = np.random.default_rng(5)
    ks = [2, 3, 17, 100]
    alphas = [np.exp(rng.uniform(np.log(0.01), np.log(10), k)).astype(np.float32) for k in ks]
    samples = [64, 32, 16, 24]
    seeds = [5, 6, 7, 8]
    urng = er.Vmt19937_64(seeds=seeds)
    res = er.dirichlet_batch(urng, alphas, samples)
    for i, k in enumerate(ks):
        r64, g = orc_mv.dirichlet(orc.Engines(seeds=[seeds[i]]), alphas[i], samples[i])
        r = res[i].cpu().numpy().astype(np.float64)
        # gamma draws are bit-exact; the float sum over K categories and the
        # division add at most (K+1) float ulps relative to the float64 result
        assert np.all(np.abs(r - r64) <= (k + 1) * 2.0 ** -23 * np.abs(r64) + 1e-38), k
        assert np.allclose(r.sum(axis=0), 1, atol=(k + 1) * 2.0 ** -22)


def test_dirichlet_mean(er):
    """test_mv.cpp:203-222: E[x_k] = alpha_k / sum(alpha) within 10%."""
    alpha = np.float32([0.5, 1.0, 2.5, 4.0])
    x = er.DirichletGen(alpha).generate(er.Vmt19937_64(11), 8192).cpu().numpy()
    assert np.allclose(x.mean(axis=1), alpha / alpha.sum(), rtol=0.1)


def test_alias_table_matches_oracle(er, orc):
    rng = np.random.default_rng(12)
    for K in (3, 15, 16, 100, 1000, 10000):
        w = rng.exponential(1.0, K)
        t = er.alias_table(w)
        if K < 16:
            c = np.cumsum(w)
            np.testing.assert_array_equal(t[1], (c / c[-1] * 0x80000000).astype(np.uint32))
        else:
            prob, alias = orc.alias_table(w)
            np.testing.assert_array_equal(t[1], prob)
            np.testing.assert_array_equal(t[2], alias)


@pytest.mark.parametrize("K", [4, 15, 16, 37, 1000])
def test_discrete_fill_matches_oracle(er, orc, K):
    rng = np.random.default_rng(K)
    w = rng.uniform(0, 1, K)
    seeds = [1, 2, 3]
    urng = er.Vmt19937_64(seeds=seeds)
    ref = orc.Engines(seeds=seeds)
    for n in (4, 8, 1004, 12, 5004):
        out = er.discrete(n, 1, urng, w)
        assert_same(out.reshape(len(seeds), n).cpu().numpy(), orc.fill_discrete_i32(ref, n, w))
    np.testing.assert_array_equal(urng.snapshot(), ref.raw)


@pytest.mark.parametrize("K", [3, 15, 16, 37, 1000, 10000])
def test_discrete_f_fill_matches_oracle(er, orc, K):
    """discreteF (float tables): one half draw per packet for K < 16, two
    otherwise; odd packet counts leave cached halves across fills."""
    rng = np.random.default_rng(K + 1)
    w = rng.uniform(0, 1, K)
    seeds = [1, 2, 3]
    urng = er.Vmt19937_64(seeds=seeds)
    ref = orc.Engines(seeds=seeds)
    for n in (4, 12, 1004, 8, 4):
        out = er.discrete(n, 1, urng, w, precision="float")
        assert_same(out.reshape(len(seeds), n).cpu().numpy(), orc.fill_discrete_f_i32(ref, n, w))
        np.testing.assert_array_equal(urng.snapshot(), ref.raw)


@pytest.mark.parametrize("cases", [
    [(20, 4, 16), (99, 10, 24), (7, 3, 8), (100, 1, 8)],           # per-sample discrete draws
    [(1000, 4, 64), (5000, 100, 16), (3001, 3, 8)],                # BinomialVGen<float> rows
    [(50, 1000, 8), (1, 20, 16), (250, 12, 40), (2000, 16, 24)],  # mixed
    [(60000, 1500, 16), (5000, 2, 8)],                              # long row chains
    [(3000, 50, 72), (10000, 30, 200), (400, 9, 64)],               # more than 64 samples per engine
])
def test_multinomial_batch_matches_oracle(er, orc, cases):
    rng = np.random.default_rng(len(cases))
    trials = [c[0] for c in cases]
    weights = [rng.uniform(0.05, 1, c[1]) for c in cases]
    samples = [c[2] for c in cases]
    seeds = [10 + 4 * i for i in range(len(cases))]
    urng = er.Vmt19937_64(seeds=seeds)
    got = er.multinomial_batch(urng, trials, weights, samples)
    exp = orc.multinomial(orc.Engines(seeds=seeds), trials, weights, samples)
    for g, x, t in zip(got, exp, trials):
        g = g.cpu().numpy()
        np.testing.assert_array_equal(g, x)
        assert np.all(g.sum(axis=0) == t)


def test_multinomial_rejects_nonfinite(er):
    urng = er.Vmt19937_64(1)
    with pytest.raises(er.ErdError):
        er.multinomial_batch(urng, [10], [[1.0, np.inf, 2.0]], [8])


def test_factors_once_generate_many(er):
    """MvNormalBatch / WishartBatch factorise at construction; each generate()
    equals the one-shot batch call on the same engine states."""
    rng = np.random.default_rng(77)
    dims = [5, 24, 130]
    covs = [rand_cov(rng, d) for d in dims]
    means = [rng.normal(size=d).astype(np.float32) for d in dims]
    seeds = [7, 11, 13]
    gen = er.MvNormalBatch(covs, means)
    a = er.Vmt19937_64(seeds=seeds)
    b = er.Vmt19937_64(seeds=seeds)
    for n in (16, 40):
        xs = gen.generate(a, n)
        ys = er.mvnormal_batch(b, covs, n, means)
        for x, y in zip(xs, ys):
            assert_same(x.cpu().numpy(), y.cpu().numpy())
    np.testing.assert_array_equal(a.snapshot(), b.snapshot())
    wgen = er.WishartBatch(covs[:2], [d + 3 for d in dims[:2]])
    w1 = wgen.generate(a, 8)
    w2 = er.wishart_batch(b, covs[:2], [d + 3 for d in dims[:2]], 8)
    for x, y in zip(w1, w2):
        assert_same(x.cpu().numpy(), y.cpu().numpy())


def test_factors_from_device_resident_covariances(er):
    """A flat float64 CUDA tensor plus dims gives the same factors (and so the
    same samples) as the list of host matrices."""
    import torch
    rng = np.random.default_rng(78)
    dims = [3, 33, 70]
    covs = [rand_cov(rng, d) for d in dims]
    flat = torch.from_numpy(np.concatenate([c.reshape(-1, order="F") for c in covs])).cuda()
    a = er.Vmt19937_64(seeds=[1, 2, 3])
    b = er.Vmt19937_64(seeds=[1, 2, 3])
    for x, y in zip(er.MvNormalBatch(flat, dims=dims).generate(a, 24), er.MvNormalBatch(covs).generate(b, 24)):
        assert_same(x.cpu().numpy(), y.cpu().numpy())
    for inverse in (False, True):
        w1 = er.WishartBatch(flat, [d + 2 for d in dims], inverse=inverse, dims=dims).generate(a, 8)
        w2 = er.WishartBatch(covs, [d + 2 for d in dims], inverse=inverse).generate(b, 8)
        for x, y in zip(w1, w2):
            assert_same(x.cpu().numpy(), y.cpu().numpy())
    with pytest.raises(ValueError):
        er.MvNormalBatch(flat[:-1], dims=dims)


def test_factors_report_not_psd(er):
    bad = -np.eye(6)
    with pytest.raises(er.ErdError):
        er.MvFactors([np.eye(3), bad])


@pytest.mark.parametrize("K", [3, 15, 16, 37, 1000, 10000])
def test_discrete_d_fill_matches_oracle(er, orc, K):
    """discreteD (double tables): K < 16 one full draw per packet; otherwise
    half() index bits and one full draw of doubles per packet. Fills of odd
    packet counts leave the half cache set, also for the next fill; one fill
    starts on a cache left by a discreteF fill."""
    rng = np.random.default_rng(K + 2)
    w = rng.uniform(0, 1, K)
    seeds = [1, 2, 3]
    urng = er.Vmt19937_64(seeds=seeds)
    ref = orc.Engines(seeds=seeds)
    for n in (4, 12, 1004, 8, 4, 4, 2500):
        out = er.discrete(n, 1, urng, w, precision="double")
        assert_same(out.reshape(len(seeds), n).cpu().numpy(), orc.fill_discrete_d_i32(ref, n, w))
        np.testing.assert_array_equal(urng.snapshot(), ref.raw)
    er.discrete(12, 1, urng, w, precision="float")  # three packets of half draws
    orc.fill_discrete_f_i32(ref, 12, w)
    out = er.discrete(1000, 1, urng, w, precision="double")
    assert_same(out.reshape(len(seeds), 1000).cpu().numpy(), orc.fill_discrete_d_i32(ref, 1000, w))
    np.testing.assert_array_equal(urng.snapshot(), ref.raw)


def test_dirichlet_gamma_rows_are_exact(er, orc):
    """The category rows of a Dirichlet instance run as one serial chain of
    GammaGen fills on the instance's engine, served by a four-wave workgroup
    (every regime: GS for alpha < 1, the exponential for 1, the integer
    products up to 19, tangent above, and 20 as tangent). The gamma rows must
    be the oracle's bit for bit: with the float column sums taken in category
    order, the normalised values are compared exactly, and the engine records
    must equal the oracle's after the same draws."""
    import orc_mv
    rng = np.random.default_rng(21)
    alphas = [np.float32([0.3, 1.0, 2.0, 0.05, 19.0, 20.0, 3.5, 1.0, 7.0, 0.9]),
              np.exp(rng.uniform(np.log(0.01), np.log(10), 1000)).astype(np.float32),
              np.exp(rng.uniform(np.log(0.01), np.log(10), 300)).astype(np.float32),
              np.float32([4.0, 0.5])]
    samples = [64, 1024, 40, 8]
    seeds = [31, 32, 33, 34]
    urng = er.Vmt19937_64(seeds=seeds)
    print("dirichlet rows", [len(a) for a in alphas], samples, flush=True)
    res = er.dirichlet_batch(urng, alphas, samples)
    snap = urng.snapshot().reshape(len(seeds), -1)
    for i, a in enumerate(alphas):
        ref = orc.Engines(seeds=[seeds[i]])
        _, g = orc_mv.dirichlet(ref, a, samples[i])
        s = np.zeros(samples[i], np.float32)
        for k in range(len(a)):
            s = (s + g[k]).astype(np.float32)
        want = (g / s).astype(np.float32)
        assert_same(res[i].cpu().numpy(), want)
        np.testing.assert_array_equal(snap[i], ref.raw)


def test_dirichlet_split_batches_match_oracle(er, orc):
    """A Dirichlet batch runs its longest chains of category rows (work at
    least 0.7 of the longest) and their normalisation on the caller's
    stream, and the other instances on a second stream at the same time;
    the caller's stream then waits for the second. Each instance must be the
    oracle's gamma rows divided by their category-order column sums, bit
    for bit, with its engine record advanced as the oracle's, whichever
    stream ran it: a short 2048-sample instance beside a long one, a
    2056-sample instance beside a short one, and a 2100-category instance
    beside a short one."""
    import orc_mv
    rng = np.random.default_rng(23)
    long_k = np.exp(rng.uniform(np.log(0.05), np.log(8), 40)).astype(np.float32)
    wide_k = np.exp(rng.uniform(np.log(0.05), np.log(8), 2100)).astype(np.float32)
    cases = [([np.float32([0.7, 2.0, 5.5]), long_k], [2048, 1024]),
             ([np.float32([0.7, 2.0, 5.5]), np.float32([1.0, 0.2, 3.0, 12.5, 0.05])], [2056, 16]),
             ([wide_k, np.float32([0.7, 2.0, 5.5])], [24, 16])]
    for alphas, samples in cases:
        seeds = [61, 62]
        urng = er.Vmt19937_64(seeds=seeds)
        print("dirichlet samples", samples, [len(a) for a in alphas], flush=True)
        res = er.dirichlet_batch(urng, alphas, samples)
        snap = urng.snapshot().reshape(len(seeds), -1)
        for i, a in enumerate(alphas):
            ref = orc.Engines(seeds=[seeds[i]])
            _, g = orc_mv.dirichlet(ref, a, samples[i])
            s = np.zeros(samples[i], np.float32)
            for k in range(len(a)):
                s = (s + g[k]).astype(np.float32)
            assert_same(res[i].cpu().numpy(), (g / s).astype(np.float32))
            np.testing.assert_array_equal(snap[i], ref.raw)
