"""FP64 oracle of the multivariate generators (test infrastructure only).

Restates MvNormalGen / WishartGen / DirichletGen (EigenRand/MvDists/
MvNormal.h, Multinomial.h) on top of the C oracle's univariate fills, with
the factorisation and products in float64 (numpy/LAPACK). The univariate
draws are bit-exact to the reference's; the float64 products are the
reference point against which the GPU's float results are bounded.
"""
import numpy as np

import pyoracle as orc


def cholesky(a):
    """detail::get_lt (MvNormal.h:20-38) in float64 (LLT branch)."""
    return np.linalg.cholesky(np.asarray(a, dtype=np.float64))


def mvnormal(eng, cov, samples, mean=None):
    """MvNormalGen::generate (MvNormal.h:103-108): (L * Z).colwise() + mean,
    Z = StdNormalGen fill of d x samples (column-major). Returns (X64, Z)."""
    d = cov.shape[0]
    z = orc.fill_normal_f32(eng, d * samples).reshape(samples, d).T
    x = cholesky(cov) @ z.astype(np.float64)
    if mean is not None:
        x = x + np.asarray(mean, dtype=np.float64)[:, None]
    return x, z


def wishart(eng, df, scale, samples):
    """WishartGen::generate (MvNormal.h:236-277) with float64 products.
    Returns (W64 as d x d*samples, A as d x d*samples float32)."""
    d = scale.shape[0]
    nrm = orc.fill_normal_f32(eng, samples * d * (d - 1) // 2)
    rand_mat = np.zeros((d, d * samples), dtype=np.float32)
    ptr = 0
    for j in range(samples):
        for i in range(d - 1):
            rand_mat[i + 1:, i + j * d] = nrm[ptr:ptr + d - 1 - i]
            ptr += d - 1 - i
    for i in range(d):
        chi = orc.fill_gamma_f32(eng, samples, np.float32(df - i) * np.float32(0.5), 2.0)
        rand_mat[i, i + np.arange(samples) * d] = np.sqrt(chi)
    t = cholesky(scale) @ rand_mat.astype(np.float64)
    w = np.empty((d, d * samples))
    for j in range(samples):
        tj = t[:, j * d:(j + 1) * d]
        w[:, j * d:(j + 1) * d] = tj @ tj.T
    return w, rand_mat


def invwishart(eng, df, scale, samples):
    """InvWishartGen::generate (MvNormal.h:404-440) with float64 products:
    the same random matrices as WishartGen, L = cholesky(inv(scale)),
    T = L A, U = T^-1, W = U^T U. Returns (W64, A)."""
    d = scale.shape[0]
    _, a = wishart(eng, df, np.eye(d), samples)
    t = cholesky(np.linalg.inv(scale)) @ a.astype(np.float64)
    w = np.empty((d, d * samples))
    for j in range(samples):
        u = np.linalg.inv(t[:, j * d:(j + 1) * d])
        w[:, j * d:(j + 1) * d] = u.T @ u
    return w, a


def dirichlet(eng, alpha, samples):
    """DirichletGen::generate (Multinomial.h:190-203): one gamma fill per
    category, then column normalisation (sum in float64). Returns (ret64, G)."""
    g = np.stack([orc.fill_gamma_f32(eng, samples, float(a), 1.0) for a in alpha])
    return g.astype(np.float64) / g.astype(np.float64).sum(axis=0, keepdims=True), g
