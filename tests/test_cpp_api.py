"""The C++ drop-in header (include/fm/sparse_matrix.h) compiles against the
C ABI (CPU) and drives the HIP path like the reference's drivers (GPU)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "test_fm_api.cpp")
BIN = os.path.join(ROOT, "tests", "cpp", "test_fm_api")
LIBDIR = os.path.join(ROOT, "flashx_amd", "lib")


def compile_driver():
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), SRC,
           "-L" + LIBDIR, "-lfmx", "-Wl,-rpath," + LIBDIR, "-o", BIN]
    subprocess.check_call(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.STDOUT)


def test_cpp_api_compiles():
    if not os.path.exists(os.path.join(LIBDIR, "libfmx.so")):
        pytest.skip("libfmx.so not built")
    compile_driver()
    assert os.path.exists(BIN)


@pytest.mark.gpu
def test_cpp_api_runs():
    if not os.path.exists(BIN):
        compile_driver()
    out = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "OK" in out.stdout
