"""The C++ face of the boundary (include/nsfem_gpu_device.hpp, the
reference's GPUSimulationDevice interface) driven by a compiled C++ caller:
triplet upload, products, BLAS-1, CG with the reference's absolute eps and
advection, each checked against host closed forms (tests/cpp)."""
import os
import subprocess

import pytest

from conftest import ROOT

BIN = os.path.join(ROOT, "tests", "cpp", "test_device_shim")


@pytest.mark.gpu
def test_cpp_device_shim():
    assert os.path.exists(BIN), "tests/cpp/test_device_shim is not built (run build())"
    out = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "PASSED" in out.stdout
