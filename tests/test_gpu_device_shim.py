"""The C++ face of the boundary (include/nsfem_gpu_device.hpp, the
reference's GPUSimulationDevice interface) driven by a compiled C++ caller:
triplet upload, products, BLAS-1, CG with the reference's absolute eps and
advection, each checked against host closed forms (tests/cpp).  The C++ test
is loaded into this process as a shared library: a second process could not
launch cooperative kernels once this one has."""
import ctypes as C
import os

import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "tests", "cpp", "libtest_device_shim.so")


@pytest.mark.gpu
def test_cpp_device_shim(capfd):
    from navier_stokes_fem_amd import _native

    _native.lib()  # the product library, loaded first: the test links against it
    assert os.path.exists(LIB), "tests/cpp/libtest_device_shim.so is not built (run build())"
    lib = C.CDLL(LIB)
    lib.nsfem_device_shim_selftest.restype = C.c_int
    failures = lib.nsfem_device_shim_selftest()
    out = capfd.readouterr().out
    print(out)
    assert failures == 0, out
    assert "PASSED" in out
