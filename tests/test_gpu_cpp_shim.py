"""The C++ front end (include/eigenrand_amd.hpp) on the GPU: its example
program's results equal the same calls made through the Python binding."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "cpp", "shim_example")


def load(d, name, dtype):
    return np.fromfile(os.path.join(d, name + ".bin"), dtype=dtype)


def test_cpp_front_end_matches_python(er, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/shim_example is not built (run __graft_entry__.build())")
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    urng = er.Vmt19937_64(42)
    np.testing.assert_array_equal(load(tmp_path, "raw", np.uint64), er.randbits_u64(256, urng).cpu().numpy().ravel())
    flat = lambda t: t.t().contiguous().view(-1).cpu().numpy()
    np.testing.assert_array_equal(load(tmp_path, "uniform", np.float32), flat(er.uniform_real(32, 32, urng)))
    np.testing.assert_array_equal(load(tmp_path, "normal", np.float32), flat(er.normal(64, 64, urng)))
    np.testing.assert_array_equal(load(tmp_path, "gamma", np.float32), flat(er.gamma(128, 1, urng, 3.7, 2.0)))
    np.testing.assert_array_equal(load(tmp_path, "poisson", np.int32), flat(er.poisson(128, 1, urng, 20.0)))
    pair = er.Vmt19937_64(seeds=[5, 9])
    gen = er.Generator("gamma", 0.4, 1.5, 2)
    want = np.concatenate([gen.generate(8, pair).cpu().numpy(), gen.generate(24, pair).cpu().numpy()])
    np.testing.assert_array_equal(load(tmp_path, "gamma_gen", np.float32), want)
    p8 = er.P8_mt19937_64(seeds=[0])
    np.testing.assert_array_equal(load(tmp_path, "p8", np.uint64), er.p8_randbits_u64(64, p8).cpu().numpy().ravel())
