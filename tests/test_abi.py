"""The C-ABI library loads and exports every symbol the public header declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "eigenrand_amd.h")
LIB = os.path.join(ROOT, "eigenrand_amd", "liberd_amd.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(erd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "erd_fill_std_normal_f32" in names and "erd_engines_create" in names
    assert len(names) >= 20


def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from eigenrand_amd import _lib
    assert set(declared()) <= set(_lib.declared_symbols())


def test_state_bytes_without_gpu():
    from eigenrand_amd import _lib, STATE_BYTES
    assert _lib.load().erd_engine_state_bytes() == STATE_BYTES


def test_cpp_front_end_compiles():
    """include/eigenrand_amd.hpp builds warning-free with the host compiler."""
    import shutil
    import subprocess
    import tempfile
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    src = "#include \"eigenrand_amd.hpp\"\nint main() { return 0; }\n"
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "t.cpp")
        with open(path, "w") as f:
            f.write(src)
        r = subprocess.run([cxx, "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                            "-I", os.path.join(ROOT, "include"), path], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("std", ["c++11", "c++14", "c++17", "c++20"])
def test_eigen_front_end_compiles(std):
    """include/eigenrand_amd/EigenRand and its example program (written the
    way EigenRand code is written, over an Eigen-like array type) build
    warning-free as C++11 (the reference's standard) and later standards."""
    import shutil
    import subprocess
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    src = os.path.join(ROOT, "tests", "cpp", "eigen_frontend_example.cpp")
    r = subprocess.run([cxx, f"-std={std}", "-Wall", "-Wextra", "-pedantic", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_vop_parameter_tensors_are_checked_before_launch():
    """Host tensors and wrong dtypes raise InvalidArgument before any engine,
    allocation or launch is touched (so no GPU is needed to see it)."""
    import torch
    import eigenrand_amd as er
    f32 = torch.ones(8, dtype=torch.float32)       # host memory
    i32 = torch.ones(8, dtype=torch.int32)
    cases = [
        lambda: er.gamma_vop(None, f32, f32),
        lambda: er.poisson_vop(None, f32),
        lambda: er.normal_vop(None, f32, f32),
        lambda: er.binomial_vop(None, i32, f32),
        lambda: er.exponential_vop(None, f32),
        lambda: er.gamma_vop(None, f32.half(), f32),
        lambda: er.poisson_vop(None, f32.double()),
        lambda: er.binomial_vop(None, i32.long(), f32),
        lambda: er.binomial_vop(None, i32.float(), f32),
        lambda: er.gamma_vop(None, np.ones(8, np.float32), f32),
    ]
    for fn in cases:
        with pytest.raises(er.InvalidArgument):
            fn()
