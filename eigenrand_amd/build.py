"""Build recipe for the native library (HIP for gfx950) and the oracle.

`build_library()` compiles every .hip source under csrc/ into
eigenrand_amd/liberd_amd.so with hipcc. The flags are part of the numerical
contract: no FP contraction (the reference's AVX2 packet code has no FMA),
correctly rounded fp32 division and square root, fp32 denormals kept, and no
packed-FP32 instructions.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liberd_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


OBJ_DIR = os.path.join(HERE, "_build")


def _compile(src, verbose):
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *HIP_FLAGS, "-c", "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-o", obj, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return obj


def build_library(verbose=False, jobs=None):
    """Each source compiles to its own object (in parallel), then one link.
    No device code is shared across sources (each kernel is launched from the
    file that defines it), so no relocatable device code is needed."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(OBJ_DIR, exist_ok=True)
    jobs = jobs or min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda f: _compile(f, verbose), sources()))
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", LIB + ".tmp", *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


CXX = os.environ.get("CXX", "g++")
SHIM_EXAMPLE = os.path.join(ROOT, "tests", "cpp", "shim_example")
EIGEN_EXAMPLE = os.path.join(ROOT, "tests", "cpp", "eigen_frontend_example")


def build_shim_example(verbose=False):
    """The C++ front ends' example programs, linked against the library: the
    erd::rand header (C++17) and the Eigen::Rand header (built as C++11, the
    standard the reference requires)."""
    for exe, std in ((SHIM_EXAMPLE, "c++17"), (EIGEN_EXAMPLE, "c++11")):
        cmd = [CXX, f"-std={std}", "-O2", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"), exe + ".cpp",
               "-L", HERE, "-lerd_amd", "-Wl,-rpath," + HERE, "-Wl,-rpath,$ORIGIN/../../eigenrand_amd",
               "-o", exe]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return SHIM_EXAMPLE


def build_oracle(verbose=False):
    cmd = ["make", "-s", "-C", os.path.join(ROOT, "oracle")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


if __name__ == "__main__":
    build_oracle(verbose=True)
    build_library(verbose=True)
    build_shim_example(verbose=True)
    print("built", LIB)
