"""The C-ABI library loads and exports every symbol the public header declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

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
