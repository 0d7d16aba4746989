"""Build the HIP library libhodlr_amd.so in-tree for gfx950.

    python -m fastsolver2_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhodlr_amd.so")
SOURCES = ["hodlr_kernels.hip", "hodlr_capi.cpp", "hodlr_batch.cpp", "hodlr_pair.cpp",
           "hodlr_facade.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed-FP32 VALU forms are disabled for every HIP translation unit.
FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         # fully unroll the register-resident panel loops (static register indexing)
         "-mllvm", "-pragma-unroll-threshold=1000000",
         # MFMA accumulators in VGPRs (gfx950 allows it): the AGPR form made the
         # compiler copy the accumulators across every k-chunk in the small GEMMs
         "-mllvm", "-amdgpu-mfma-vgpr-form",
         "-Wall", "-Wno-unused-result"]


def build_facade_test(verbose: bool = False) -> str:
    """tests/cpp/test_facade (the C++ facade tests), linked against the library."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "test_facade.cpp")
    out = os.path.join(root, "tests", "cpp", "test_facade")
    if not os.path.exists(src):
        return ""
    cmd = [HIPCC, "-O2", "-std=c++17", "-I", os.path.join(root, "include"), src, "-L", HERE,
           "-lhodlr_amd", "-Wl,-rpath,$ORIGIN/../../fastsolver2_amd", "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("facade test build failed:\n" + res.stdout + res.stderr)
    os.replace(out + ".tmp", out)
    return out


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-I", os.path.join(HERE, "..", "include"), "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (res.stdout + res.stderr).splitlines()
                    if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + out)
    if verbose and out.strip():
        print(out)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_facade_test(verbose="-v" in sys.argv))
