"""Build the HIP library (gfx950) in-tree: cps_veloc_amd/libcps_dwf_gpu.so."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libcps_dwf_gpu.so")
SOURCES = [os.path.join(HERE, "csrc", "dwf_gpu.hip"), os.path.join(HERE, "csrc", "synth.cpp"),
           os.path.join(HERE, "csrc", "nersc.cpp")]
HEADERS = [os.path.join(HERE, "csrc", "dwf_kernels.hpp"),
           os.path.join(HERE, "csrc", "ptc_kernels.hpp"),
           os.path.join(ROOT, "include", "dwf_gpu.h"),
           os.path.join(ROOT, "include", "roofline.h"),
           os.path.join(ROOT, "include", "dwf_synth.h")]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc, *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(HERE, "csrc"), "-o", LIB + ".tmp", *SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
