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


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
