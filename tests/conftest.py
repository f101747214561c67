import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.build()
    return O


@pytest.fixture(scope="session")
def native():
    from navier_stokes_fem_amd import _native

    return _native.lib()


@pytest.fixture(scope="session")
def gpu():
    """One device handle for the whole GPU session."""
    from navier_stokes_fem_amd.device import GPUSimulationDevice

    dev = GPUSimulationDevice(0)
    yield dev
    dev.close()


@pytest.fixture(scope="session")
def mesh_small():
    """assets/mesh_small.json of the reference, stored as arrays."""
    from navier_stokes_fem_amd.mesh import Mesh

    z = np.load(os.path.join(GOLDEN, "mesh_small.npz"))
    return Mesh(nodes=z["nodes"], elements=z["elements"], pressure_count=int(z["pressure_count"]),
                vel_bc_nodes=z["vel_bc_nodes"], vel_bc_u=z["vel_bc_u"], vel_bc_v=z["vel_bc_v"],
                p_bc_nodes=z["p_bc_nodes"], p_bc_values=z["p_bc_values"], name="mesh_small")


def random_csr(n, m, density, rng, ensure_diag=False):
    """Random CSR with sorted unique columns per row (some empty rows)."""
    rows, cols, vals = [], [], []
    for r in range(n):
        k = rng.binomial(m, density)
        c = np.sort(rng.choice(m, size=min(k, m), replace=False))
        if ensure_diag and r < m and r not in c:
            c = np.sort(np.append(c, r))
        rows.append(np.full(len(c), r))
        cols.append(c)
        vals.append(rng.standard_normal(len(c)))
    rs = np.zeros(n + 1, np.int32)
    rs[1:] = np.cumsum([len(c) for c in cols])
    col = np.concatenate(cols).astype(np.int32) if cols else np.zeros(0, np.int32)
    val = np.concatenate(vals) if vals else np.zeros(0)
    return rs, col, val


def dense(rs, col, val, shape):
    A = np.zeros(shape)
    for r in range(shape[0]):
        A[r, col[rs[r]:rs[r + 1]]] += val[rs[r]:rs[r + 1]]
    return A
