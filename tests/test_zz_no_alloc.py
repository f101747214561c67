"""The no-allocation CG entry point (nsfem_cg_solve_noalloc) against the
allocating one (nsfem_cg_solve), called through the C ABI: the same
kernels on the same system give bit-identical solutions and statistics, on
two calls in a row (the second on a smaller system, so the scratch array
holds the first call's larger state), in FP64 and in FP32 + FP64
refinement; systems beyond the scratch capacity are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(dev, slot, m, name, nu=1e-3, dt=1e-3):
    from navier_stokes_fem_amd.device import assemble

    A = assemble(m, nu, dt)[name]
    rs, col, val, shape = A
    dev.upload_csr(slot, rs, col, val, shape=shape)
    return shape[0]


def _solve(dev, entry, slot, b, x0, columns, precision, rtol=1e-10):
    from navier_stokes_fem_amd import _native as N

    prec = {"fp64": N.PRECISION_FP64, "fp32_refined": N.PRECISION_FP32_REFINED}[precision]
    bb = dev.buffer(b)
    xb = dev.buffer(x0)
    stats = (N.CgStats * 2)()
    st = getattr(N.lib(), entry)(dev.handle, int(slot), bb.handle, xb.handle, int(columns), -1,
                                 float(rtol), 0.0, prec, stats)
    x = xb.download()
    bb.close()
    xb.close()
    return st, x, [stats[c].as_dict() for c in range(columns)]


def _same(a, b):
    sa, xa, ta = a
    sb, xb, tb = b
    assert sa == sb == 0
    np.testing.assert_array_equal(xa, xb)
    for u, v in zip(ta, tb):
        for k in ("iterations", "refinements", "status", "columns", "rhs_norm",
                  "initial_residual", "final_residual", "tolerance"):
            assert u[k] == v[k], (k, u[k], v[k])


@pytest.mark.parametrize("precision", ["fp64", "fp32_refined"])
def test_noalloc_solve_matches_the_allocating_solve(gpu, precision):
    from navier_stokes_fem_amd import mesh

    rng = np.random.default_rng(17)
    # first call: the 256^2 cavity's pressure system (66,049 rows), one column
    n1 = _system(gpu, 0, mesh.cavity(256), "pressure_stiffness")
    b1 = rng.standard_normal(n1)
    x1 = rng.standard_normal(n1) * 0.1
    ref1 = _solve(gpu, "nsfem_cg_solve", 0, b1, x1, 1, precision)
    got1 = _solve(gpu, "nsfem_cg_solve_noalloc", 0, b1, x1, 1, precision)
    _same(ref1, got1)
    # second call, smaller: the 48^2 cavity's diffusion system, two columns
    n2 = _system(gpu, 2, mesh.cavity(48), "diffusion")
    assert 2 * n2 < n1
    b2 = rng.standard_normal(2 * n2)
    x2 = np.zeros(2 * n2)
    got2 = _solve(gpu, "nsfem_cg_solve_noalloc", 2, b2, x2, 2, precision)
    ref2 = _solve(gpu, "nsfem_cg_solve", 2, b2, x2, 2, precision)
    _same(ref2, got2)
    print(f"{precision}: iterations {got1[2][0]['iterations']} (n={n1}), "
          f"{[s['iterations'] for s in got2[2]]} (n={n2}, 2 columns)")


def test_noalloc_solve_refuses_systems_beyond_its_scratch(gpu):
    from navier_stokes_fem_amd import _native as N

    n = (1 << 20) + 1
    rs = np.arange(n + 1, dtype=np.int32)
    col = np.arange(n, dtype=np.int32)
    gpu.upload_csr(1, rs, col, np.full(n, 2.0))
    st, _, _ = _solve(gpu, "nsfem_cg_solve_noalloc", 1, np.ones(n), np.zeros(n), 1, "fp64")
    assert st == N.INVALID_ARGUMENT
    assert "scratch" in N.last_error()
    # the allocating entry point solves it (diagonal: one iteration)
    st, x, stats = _solve(gpu, "nsfem_cg_solve", 1, np.ones(n), np.zeros(n), 1, "fp64")
    assert st == 0 and np.allclose(x, 0.5)
    # one column fewer fits
    n = 1 << 20
    gpu.upload_csr(1, rs[:n + 1], col[:n], np.full(n, 2.0))
    st, x, _ = _solve(gpu, "nsfem_cg_solve_noalloc", 1, np.ones(n), np.zeros(n), 1, "fp64")
    assert st == 0 and np.allclose(x, 0.5)
