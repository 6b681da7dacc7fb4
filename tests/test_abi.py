"""The C-ABI library loads and exports every symbol the public header declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

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
