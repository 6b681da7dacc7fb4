"""The Eigen::Rand front end (include/eigenrand_amd/EigenRand) on the GPU: a
program written like EigenRand code (normal<Derived>(rows, cols, urng),
xxxLike, generator objects, vectorize-over-params overloads, a P8 adaptor as
the engine, a device-resident destination, the multivariate generators) gives
the same values as the same draws made through the Python binding, whose
fills are checked against the oracle elsewhere."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "cpp", "eigen_frontend_example")


def load(d, name, dtype):
    return np.fromfile(os.path.join(d, name + ".bin"), dtype=dtype)


def col(t):
    """Column-major linear order of a (rows, cols) result."""
    return t.t().contiguous().view(-1).cpu().numpy()


def test_eigen_front_end_matches_python(er, tmp_path):
    import torch
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/eigen_frontend_example is not built (run __graft_entry__.build())")
    print("running", EXE, flush=True)
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    eq = np.testing.assert_array_equal

    urng = er.Vmt19937_64(42)
    eq(load(tmp_path, "normal", np.float32), col(er.normal(64, 32, urng)))
    eq(load(tmp_path, "normal_like", np.float32), col(er.normal(64, 32, urng, 2.0, 0.5)))
    eq(load(tmp_path, "uniform", np.float32), col(er.uniform_real(32, 32, urng, -1.0, 2.0)))
    eq(load(tmp_path, "gamma", np.float32), col(er.gamma(128, 1, urng, 3.7, 2.0)))
    eq(load(tmp_path, "poisson", np.int32), col(er.poisson(128, 1, urng, 20.0)))
    eq(load(tmp_path, "binomial", np.int32), col(er.binomial(8, 8, urng, 100, 0.3)))
    eq(load(tmp_path, "discrete", np.int32), col(er.discrete(64, 1, urng, [1.0, 2.0, 3.0, 4.0])))
    eq(load(tmp_path, "normal_f64", np.float64), col(er.normal_f64(16, 4, urng)))
    eq(load(tmp_path, "chisq", np.float32), col(er.chi_squared(64, 1, urng, 3.0)))
    mean = torch.arange(64, dtype=torch.float32, device="cuda") - 32.0
    eq(load(tmp_path, "normal_vs", np.float32), er.normal_vop(urng, mean, 2.0).cpu().numpy())
    i = np.arange(64)
    trials = torch.from_numpy((10 * i + 1).astype(np.int32)).cuda()
    p = torch.from_numpy(((i + 1) / 65.0).astype(np.float32)).cuda()
    eq(load(tmp_path, "binomial_vv", np.int32), er.binomial_vop(urng, trials, p).cpu().numpy())

    pair = er.Vmt19937_64(seeds=[5, 9])
    gen = er.Generator("gamma", 0.4, 1.5, 2)
    eq(load(tmp_path, "gamma_gen1", np.float32), gen.generate(8, pair).cpu().numpy())
    eq(load(tmp_path, "gamma_gen2", np.float32), gen.generate(24, pair).cpu().numpy())

    p8 = er.P8_mt19937_64(seeds=[42])
    eq(load(tmp_path, "p8_normal", np.float32), er.p8_fill("std_normal", 24, p8).cpu().numpy().ravel())
    eq(load(tmp_path, "p8_poisson", np.int32), er.p8_fill("poisson", 12, p8, 3.0).cpu().numpy().ravel())

    eq(load(tmp_path, "device_uniform", np.float32), col(er.uniform_real(256, 1, er.Vmt19937_64(7))))

    mvu = er.Vmt19937_64(11)
    cov = np.full((4, 4), 0.5) + 1.5 * np.eye(4)
    mu = np.arange(4, dtype=np.float32)
    eq(load(tmp_path, "mvnormal", np.float32), col(er.mvnormal_batch(mvu, [cov], 64, means=[mu])[0]))
    eq(load(tmp_path, "wishart", np.float32), col(er.wishart_batch(mvu, [cov], 6, 8)[0]))
    w = np.arange(1, 6, dtype=np.float64)
    eq(load(tmp_path, "multinomial", np.int32), col(er.multinomial_batch(mvu, 20, [w], 16)[0]))
    eq(load(tmp_path, "dirichlet", np.float32), col(er.dirichlet_batch(mvu, [w.astype(np.float32)], 16)[0]))
