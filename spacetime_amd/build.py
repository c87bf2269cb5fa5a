"""Build libgpe.so (HIP kernels + C-ABI host runtime) in-tree for gfx950 with hipcc.

    python -m spacetime_amd.build      # or: python spacetime_amd/build.py

The library lands at spacetime_amd/libgpe.so; nothing is installed elsewhere.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libgpe.so")

SOURCES = ["gpe_kernels.hip", "gpe_adjoint.hip", "gpe_host.cpp", "gpe_states.cpp", "gpe_io.cpp", "gpe_noalloc.hip"]
ARCH = os.environ.get("GPE_OFFLOAD_ARCH", "gfx950")

# Packed-FP32 VALU forms are disabled for all device code.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-I" + CSRC, "-I" + INCLUDE]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found")
    return path


def build(verbose=False):
    """Compile every source and link libgpe.so; raises on failure."""
    cc = hipcc()
    objs = []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
        cmd = [cc] + COMMON + DEVICE + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [cc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_driver(verbose=False):
    """Compile the C++ host driver of the tests (tests/cxx/gpe_driver.cpp: the C ABI used from C++
    alone) against include/ and link it to the in-tree libgpe.so; returns the binary's path."""
    src = os.path.join(ROOT, "tests", "cxx", "gpe_driver.cpp")
    out = os.path.join(ROOT, "tests", "cxx", "gpe_driver")
    cxx = shutil.which("g++") or "g++"
    cmd = [cxx, "-std=c++17", "-O2", "-Wall", "-I" + INCLUDE, src, "-L" + HERE, "-lgpe",
           "-Wl,-rpath,$ORIGIN/../../spacetime_amd", "-o", out]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_driver(verbose="-v" in sys.argv))
