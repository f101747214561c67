"""CPU test of the host-side dense SPD inverse used for the deepest
multigrid level (navier_stokes_fem_amd/csrc/spd_inverse.hpp): compiled with
g++ into a small driver, checked against numpy's inverse on a Galerkin
coarse matrix of the cavity pressure operator and on a random SPD matrix,
and shown to refuse an indefinite matrix."""
import os
import shutil
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

from navier_stokes_fem_amd import mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "navier_stokes_fem_amd", "csrc")

DRIVER = r"""
#include <cstdio>
#include <vector>
#include "spd_inverse.hpp"
int main(int argc, char **argv) {
    FILE *f = std::fopen(argv[1], "rb");
    int n = 0; long long nnz = 0;
    if (std::fread(&n, sizeof n, 1, f) != 1 || std::fread(&nnz, sizeof nnz, 1, f) != 1) return 2;
    std::vector<int> rs(n + 1), col(nnz);
    std::vector<double> val(nnz), inv;
    if (std::fread(rs.data(), sizeof(int), n + 1, f) != (size_t)(n + 1)) return 2;
    if (std::fread(col.data(), sizeof(int), nnz, f) != (size_t)nnz) return 2;
    if (std::fread(val.data(), sizeof(double), nnz, f) != (size_t)nnz) return 2;
    std::fclose(f);
    const bool ok = nsfem::spd_inverse(n, rs.data(), col.data(), val.data(), inv);
    FILE *o = std::fopen(argv[2], "wb");
    int flag = ok ? 1 : 0;
    std::fwrite(&flag, sizeof flag, 1, o);
    if (ok) std::fwrite(inv.data(), sizeof(double), inv.size(), o);
    std::fclose(o);
    return 0;
}
"""


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("spd")
    src = d / "drv.cpp"
    src.write_text(DRIVER)
    exe = d / "drv"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, str(src), "-o", str(exe)], check=True)
    return d, exe


def _run(driver, A):
    d, exe = driver
    A = sp.csr_matrix(A)
    A.sort_indices()
    inp, out = d / "in.bin", d / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.int32(A.shape[0]).tobytes())
        f.write(np.int64(A.nnz).tobytes())
        f.write(A.indptr.astype(np.int32).tobytes())
        f.write(A.indices.astype(np.int32).tobytes())
        f.write(A.data.astype(np.float64).tobytes())
    subprocess.run([str(exe), str(inp), str(out)], check=True)
    raw = np.fromfile(out, dtype=np.uint8)
    ok = int(np.frombuffer(raw[:4].tobytes(), np.int32)[0])
    if not ok:
        return None
    n = A.shape[0]
    return np.frombuffer(raw[4:].tobytes(), np.float64).reshape(n, n)


def test_inverse_of_a_galerkin_coarse_pressure_matrix(driver):
    from oracle import oracle as O

    m = mesh.cavity(32)
    asm = O.Assembly(m, 1e-3, 1e-3)
    rs, col, val = asm.csr["pressure_stiffness"]
    A = sp.csr_matrix((val, col, rs), shape=(len(rs) - 1,) * 2)
    prs, pcol, pval, nc = mesh.cavity_prolongation(32, 4)  # 9 x 9 coarse nodes
    P = sp.csr_matrix((pval, pcol, prs), shape=(A.shape[0], nc))
    Ac = (P.T @ A @ P).tocsr()
    inv = _run(driver, Ac)
    assert inv is not None
    ref = np.linalg.inv(Ac.toarray())
    np.testing.assert_allclose(inv, ref, rtol=0, atol=1e-9 * np.abs(ref).max())
    np.testing.assert_allclose(inv @ Ac.toarray(), np.eye(nc), rtol=0, atol=1e-9)


def test_inverse_of_a_random_spd_matrix(driver):
    rng = np.random.default_rng(7)
    B = rng.standard_normal((40, 40))
    A = B @ B.T + 40 * np.eye(40)
    inv = _run(driver, A)
    np.testing.assert_allclose(inv, np.linalg.inv(A), rtol=0, atol=1e-12)


def test_indefinite_matrix_is_refused(driver):
    A = np.diag([2.0, 1.0, -1.0, 3.0])
    assert _run(driver, A) is None
