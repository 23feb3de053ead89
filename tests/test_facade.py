"""The C++ facade (include/dwf_gpu.hpp) compiled the way a CPS build would
use it: g++ against include/ and the built library.  On the CPU it must
compile and link; on the GPU tests/cpp/facade_check.cpp runs its checks
(facade == C ABI bit for bit, true residual, MAXITER count, error mapping)."""
import os
import subprocess

import pytest

from cps_veloc_amd import dwf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "facade_check.cpp")


def build_driver(out_dir):
    exe = os.path.join(out_dir, "facade_check")
    libdir = os.path.dirname(dwf.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror",
                    "-I", os.path.join(ROOT, "include"), SRC, "-o", exe,
                    "-L", libdir, "-lcps_dwf_gpu", f"-Wl,-rpath,{libdir}"],
                   check=True, capture_output=True, text=True)
    return exe


@pytest.mark.skipif(not os.path.exists(dwf.LIB_PATH), reason="library not built")
def test_facade_compiles_and_links(tmp_path):
    exe = build_driver(str(tmp_path))
    assert os.path.exists(exe)


@pytest.mark.gpu
def test_facade_runs(tmp_path):
    print("[gpu] C++ facade check", flush=True)
    exe = build_driver(str(tmp_path))
    r = subprocess.run([exe], capture_output=True, text=True, timeout=100)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "facade_check OK" in r.stdout
