"""Build the engine's shared library, ``autopas_amd/libljgpu.so``, for gfx950.

Every HIP translation unit is compiled with ``hipcc --offload-arch=gfx950`` and linked into one
C-ABI library (include/ljgpu.h).  Translation units whose FP64 arithmetic must round exactly like the
reference expressions (binning, tower/cluster build, integrator) are compiled with
``-ffp-contract=off``; the force kernels keep hipcc's default contraction.  Packed FP32 instructions are
disabled for all device code.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libljgpu.so")
OBJDIR = os.path.join(HERE, "_build")

ARCH = os.environ.get("LJG_OFFLOAD_ARCH", "gfx950")
COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# translation unit -> extra flags
SOURCES = {
    "ljg_api.hip": [],
    "ljg_bin.hip": ["-ffp-contract=off"],
    "ljg_vcl.hip": ["-ffp-contract=off"],
    "ljg_md.hip": ["-ffp-contract=off"],
    "ljg_force.hip": [],
    "ljg_verlet.hip": ["-ffp-contract=off"],
    "ljg_ensemble.hip": ["-ffp-contract=off"],
    "ljg_generate.cpp": ["-ffp-contract=off"],
    "ljg_checkpoint.cpp": [],
    "ljg_noalloc.hip": ["-ffp-contract=off"],
}


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src, extra):
    obj = os.path.join(OBJDIR, src + ".o")
    cmd = [_hipcc()] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [_hipcc(), "-x", "hip"] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


ROOT = os.path.dirname(HERE)
CPP_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "autopas_api_test.cpp")
CPP_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "autopas_api_test")


def build_cpp_test():
    """Compile the driver of the C++ host API tests (include/autopas_gpu.hpp) against libljgpu.so, with the
    host compiler and no FP contraction (its integrator rounds like TimeDiscretization.cpp)."""
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
           CPP_TEST_SRC, "-L", HERE, "-lljgpu", "-Wl,-rpath,$ORIGIN/../../autopas_amd", "-o", CPP_TEST_BIN]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return CPP_TEST_BIN


def build(verbose=False):
    """Compile all sources and link libljgpu.so (always rebuilds), then the C++ API test driver."""
    os.makedirs(OBJDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv), SOURCES.items()))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_cpp_test()
    if verbose:
        print(f"built {LIB} and {CPP_TEST_BIN}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
