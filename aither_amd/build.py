"""Build the native library (HIP kernels + C ABI) for gfx950, in-tree.

    python -m aither_amd.build

writes aither_amd/libaither_amd.so.  The oracle (test infrastructure) is
built separately by oracle/Makefile.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libaither_amd.so")
SOURCES = ["host_case.cpp", "io.cpp", "solver.hip"]
ARCH = os.environ.get("AITHER_AMD_ARCH", "gfx950")

# packed FP32 instructions are disabled for all device code.  Multiply-add
# contraction is off: every product and sum is rounded separately, in the
# order the reference writes it, so the FP64 path reproduces the reference's
# arithmetic operation by operation.
HIPCC_FLAGS = [
    "-ffp-contract=off",
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def build_library(verbose=True):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    cmd += [os.path.join(CSRC, s) for s in SOURCES] + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_oracle(verbose=True):
    odir = os.path.join(ROOT, "oracle")
    cmd = ["make", "-s", "-C", odir, "liboracle.so", "liboracle_omp.so"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return os.path.join(odir, "liboracle.so")


if __name__ == "__main__":
    build_library()
    if "--oracle" in sys.argv:
        build_oracle()
