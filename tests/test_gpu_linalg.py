"""GPU parity of the CSR products, BLAS-1 and CG against the CPU oracle.

Bars: the CSR products add each row's products in row order exactly like the
reference's row loop (and the oracle), so they are compared BIT-EXACTLY.
Dot products are deterministic but tree-ordered: relative 1e-13.  CG: every
solve meets its stated tolerance with the true residual recomputed in FP64,
and the solution matches the oracle's FP64 CG within cond(A)*tol.
"""
import numpy as np
import pytest

from conftest import dense, random_csr

pytestmark = pytest.mark.gpu


def _buf(gpu, a):
    return gpu.buffer(np.ascontiguousarray(a, dtype=np.float64))


@pytest.mark.parametrize("n,m,density", [(1, 1, 1.0), (7, 5, 0.5), (300, 300, 0.02),
                                         (5000, 4000, 0.003), (2000, 300, 0.3)])
def test_spmv_bit_exact(gpu, oracle, n, m, density):
    rng = np.random.default_rng(n * 31 + m)
    rs, col, val = random_csr(n, m, density, rng)
    gpu.upload_csr(7, rs, col, val, shape=(n, m))
    x = rng.standard_normal(m)
    lhs = rng.standard_normal(n)
    bx, by, bl = _buf(gpu, x), _buf(gpu, np.zeros(n)), _buf(gpu, lhs)
    gpu.sp_r_mult(7, bx, by)
    y = by.download()
    ref = oracle.spmv((rs, col, val), x)
    assert np.array_equal(y, ref)
    assert np.allclose(y, dense(rs, col, val, (n, m)) @ x, rtol=1e-12, atol=1e-12)
    gpu.sp_r_mult_sub(7, bl, bx, bl)  # res aliases lhs (allowed by the reference)
    assert np.array_equal(bl.download(), oracle.spmv_sub((rs, col, val), lhs, x))


def test_spmv_long_row(gpu, oracle):
    """A row longer than one LDS tile takes the block-reduction path."""
    rng = np.random.default_rng(3)
    n, m = 4, 6000
    rs = np.array([0, 3, 3 + 5000, 5003 + 2, 5005 + 100], np.int32)
    col = np.concatenate([np.sort(rng.choice(m, 3, replace=False)),
                          np.sort(rng.choice(m, 5000, replace=False)),
                          np.sort(rng.choice(m, 2, replace=False)),
                          np.sort(rng.choice(m, 100, replace=False))]).astype(np.int32)
    val = rng.standard_normal(len(col))
    gpu.upload_csr(8, rs, col, val, shape=(n, m))
    x = rng.standard_normal(m)
    by = _buf(gpu, np.zeros(n))
    gpu.sp_r_mult(8, _buf(gpu, x), by)
    y = by.download()
    ref = oracle.spmv((rs, col, val), x)
    assert np.array_equal(y[[0, 2, 3]], ref[[0, 2, 3]])
    assert abs(y[1] - ref[1]) <= 1e-12 * np.abs(val[3:5003] * x[col[3:5003]]).sum()


def test_spmv_offsets_at_the_ends_and_timing(gpu, oracle):
    """A square matrix with columns r-300, r-7, r, r+7, r+300 (16-bit offsets
    from each tile's first row): the first and last groups of a tile hold
    entries of the neighbouring tiles, which decode below 0 in the first
    tile and past n-1 in the last and must be clamped, never summed.  The
    product is bit-exact; the event-timed entry point gives the same y and a
    positive time, and rejects a non-positive repeat count."""
    from navier_stokes_fem_amd import _native as N

    rng = np.random.default_rng(12)
    n = 1999
    rows = [np.unique(np.clip(r + np.array([-300, -7, 0, 7, 300]), 0, n - 1)) for r in range(n)]
    rs = np.concatenate([[0], np.cumsum([len(c) for c in rows])]).astype(np.int32)
    col = np.concatenate(rows).astype(np.int32)
    val = rng.standard_normal(len(col))
    gpu.upload_csr(8, rs, col, val)
    x = rng.standard_normal(n)
    ref = oracle.spmv((rs, col, val), x)
    bx, by = _buf(gpu, x), _buf(gpu, np.zeros(n))
    gpu.sp_r_mult(8, bx, by)
    assert np.array_equal(by.download(), ref)
    by2 = _buf(gpu, np.zeros(n))
    sec = gpu.time_sp_r_mult(8, bx, by2, repeats=5)
    assert 0.0 < sec < 1.0
    assert np.array_equal(by2.download(), ref)
    with pytest.raises(N.NsfemError):
        gpu.time_sp_r_mult(8, bx, by2, repeats=0)


def test_spmv_square_long_row(gpu, oracle):
    """A square banded matrix with one row longer than a tile: the other rows
    bit-exact, the long row (block-reduction path) within rounding."""
    rng = np.random.default_rng(4)
    n = 3000
    rows = []
    for r in range(n):
        if r == 1500:
            c = np.sort(rng.choice(np.arange(max(0, r - 1400), min(n, r + 1400)), 900, replace=False))
        else:
            c = np.unique(np.clip(r + np.array([-40, -1, 0, 1, 40]), 0, n - 1))
        rows.append(c)
    rs = np.concatenate([[0], np.cumsum([len(c) for c in rows])]).astype(np.int32)
    col = np.concatenate(rows).astype(np.int32)
    val = rng.standard_normal(len(col))
    gpu.upload_csr(8, rs, col, val)
    x = rng.standard_normal(n)
    lhs = rng.standard_normal(n)
    by, bl = _buf(gpu, np.zeros(n)), _buf(gpu, lhs)
    gpu.sp_r_mult(8, _buf(gpu, x), by)
    y = by.download()
    ref = oracle.spmv((rs, col, val), x)
    others = np.arange(n) != 1500
    assert np.array_equal(y[others], ref[others])
    k0, k1 = rs[1500], rs[1501]
    assert abs(y[1500] - ref[1500]) <= 1e-12 * np.abs(val[k0:k1] * x[col[k0:k1]]).sum()
    gpu.sp_r_mult_sub(8, bl, _buf(gpu, x), bl)
    got = bl.download()
    want = oracle.spmv_sub((rs, col, val), lhs, x)
    assert np.array_equal(got[others], want[others])


def test_triplet_upload_sums_duplicates(gpu, oracle):
    rng = np.random.default_rng(5)
    n = 50
    ti = rng.integers(0, n, 600).astype(np.int32)
    tj = rng.integers(0, n, 600).astype(np.int32)
    tv = rng.standard_normal(600)
    gpu.upload_matrix(9, n, n, ti, tj, tv)
    rs, col, val = oracle.triplet_to_csr(n, ti, tj, tv)
    x = rng.standard_normal(n)
    by = _buf(gpu, np.zeros(n))
    gpu.sp_r_mult(9, _buf(gpu, x), by)
    assert np.array_equal(by.download(), oracle.spmv((rs, col, val), x))
    assert gpu.matrix_info(9) == (n, n, len(col))


def test_blas1(gpu, oracle):
    rng = np.random.default_rng(11)
    n = 100_003
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    bx, by, bo = _buf(gpu, x), _buf(gpu, y), _buf(gpu, np.zeros(n))
    gpu.saxpy(n, 0.75, bx, by)
    assert np.allclose(by.download(), y + 0.75 * x, rtol=1e-15, atol=1e-15)
    gpu.saxpby(n, 2.0, -0.5, bx, by, bo)
    assert np.allclose(bo.download(), 2.0 * x - 0.5 * (y + 0.75 * x), rtol=1e-14, atol=1e-14)
    res = _buf(gpu, np.array([1.5]))  # dotProduct ADDS to result[0]
    gpu.dot_product(n, bx, bo, res)
    d = res.download()[0]
    ref = 1.5 + oracle.dot(x, bo.download())
    assert abs(d - ref) <= 1e-13 * (1.5 + np.abs(x * bo.download()).sum())
    res2 = _buf(gpu, np.array([1.5]))
    gpu.dot_product(n, bx, bo, res2)
    assert res2.download()[0] == d  # deterministic


def _cavity_system(oracle, n=24, which="pressure_stiffness"):
    from navier_stokes_fem_amd import mesh

    m = mesh.cavity(n)
    asm = oracle.Assembly(m, 0.01, 0.01)
    return m, asm.csr[which]


@pytest.mark.parametrize("which", ["pressure_stiffness", "diffusion", "velocity_mass"])
@pytest.mark.parametrize("precision", ["fp64", "fp32_refined"])
def test_cg_matches_oracle(gpu, oracle, which, precision):
    m, A = _cavity_system(oracle, 24, which)
    rs, col, val = A
    n = len(rs) - 1
    rng = np.random.default_rng(17)
    b = rng.standard_normal(n)
    x0 = rng.standard_normal(n) * 0.1
    gpu.upload_csr(10, rs, col, val)
    bx = _buf(gpu, x0)
    st = gpu.cg_solve(10, _buf(gpu, b), bx, rtol=1e-10, precision=precision)[0]
    x = bx.download()
    ox, ost, ores = oracle.cg(A, b, x0, rtol=1e-10)
    assert ost == 0 and st["status"] == 0
    true_res = np.linalg.norm(oracle.spmv_sub(A, b, x))
    assert true_res <= 1e-10 * np.linalg.norm(b) * (1 + 1e-9)
    assert abs(st["final_residual"] - true_res) <= 1e-3 * true_res + 1e-300
    Ad = dense(rs, col, val, (n, n))
    xd = np.linalg.solve(Ad, b)
    cond = np.linalg.cond(Ad)
    assert np.linalg.norm(x - xd) <= 10 * cond * 1e-10 * np.linalg.norm(xd)
    assert np.linalg.norm(x - ox) <= 10 * cond * 1e-10 * np.linalg.norm(ox)
    if precision == "fp64":
        assert abs(st["iterations"] - ores.iterations) <= 2
    else:
        assert st["refinements"] >= 1


# two-column mixed forms: the single-reduction iteration (narrow-gather by
# default, packed) in both block sizes, and the two-phase iteration
PAIRED_FORMS = [("fp64", {}), ("fp32_refined", {}),
                ("fp32_refined", {"NSFEM_CG_P2_BLOCK": "768", "NSFEM_CG_P2_LOADS": "4"}),
                ("fp32_refined", {"NSFEM_CG_PAIRED": "packed"}),
                ("fp32_refined", {"NSFEM_CG_PAIRED": "twophase"})]


def test_cg_two_columns(gpu, oracle, monkeypatch):
    m, A = _cavity_system(oracle, 20, "diffusion")
    rs, col, val = A
    n = len(rs) - 1
    rng = np.random.default_rng(23)
    b = rng.standard_normal(2 * n)
    b[n:] *= 3.0
    gpu.upload_csr(11, rs, col, val)
    for precision, env in PAIRED_FORMS:
        print(f"two columns {precision} {env}", flush=True)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        bx = _buf(gpu, np.zeros(2 * n))
        st = gpu.cg_solve(11, _buf(gpu, b), bx, columns=2, rtol=1e-10, precision=precision)
        x = bx.download()
        for c in range(2):
            bc = b[c * n:(c + 1) * n]
            xc = x[c * n:(c + 1) * n]
            r = np.linalg.norm(oracle.spmv_sub(A, bc, xc))
            assert st[c]["status"] == 0
            assert r <= 1e-10 * np.linalg.norm(bc) * (1 + 1e-9)
            if precision == "fp64":
                single = _buf(gpu, np.zeros(n))
                s1 = gpu.cg_solve(11, _buf(gpu, bc), single, rtol=1e-10)[0]
                assert abs(s1["iterations"] - st[c]["iterations"]) <= 1
            else:
                # within a few iterations of the one-column mixed solve
                single = _buf(gpu, np.zeros(n))
                s1 = gpu.cg_solve(11, _buf(gpu, bc), single, rtol=1e-10,
                                  precision=precision)[0]
                assert abs(s1["iterations"] - st[c]["iterations"]) <= 0.1 * s1["iterations"] + 3
        for k in env:
            monkeypatch.delenv(k)


@pytest.mark.parametrize("env", [{}, {"NSFEM_CG_P2_BLOCK": "768"}, {"NSFEM_CG_PAIRED": "packed"},
                                 {"NSFEM_CG_PAIRED": "packed", "NSFEM_CG_P2_BLOCK": "768"},
                                 {"NSFEM_CG_PAIRED": "twophase"}],
                         ids=["narrow1024", "narrow768", "packed1024", "packed768", "twophase"])
def test_cg_two_columns_one_converged(gpu, oracle, monkeypatch, env):
    """A pair whose second column starts at its solution (zero right-hand
    side, zero guess) and one whose second column converges in a handful of
    iterations: the first column still reaches the tolerance and the second
    is left exactly / within tolerance, on a system large enough for many
    persistent blocks."""
    from navier_stokes_fem_amd import device as D
    from navier_stokes_fem_amd import mesh

    for k, v in env.items():
        monkeypatch.setenv(k, v)
    m = mesh.cavity(96)
    a = D.assemble(m, 0.01, 0.01)
    rs, col, val, _ = a["velocity_mass"]
    A = (rs, col, val)
    n = len(rs) - 1
    gpu.upload_csr(15, rs, col, val)
    rng = np.random.default_rng(41)
    b0 = rng.standard_normal(n)
    e = np.zeros(n)
    e[n // 2] = 1.0
    for second, name in ((np.zeros(n), "zero"), (oracle.spmv_sub(A, np.zeros(n), -e), "one node")):
        print(f"two columns, second {name}, {env}", flush=True)
        b = np.concatenate([b0, second])
        bx = _buf(gpu, np.zeros(2 * n))
        st = gpu.cg_solve(15, _buf(gpu, b), bx, columns=2, rtol=1e-10, precision="fp32_refined")
        x = bx.download()
        assert st[0]["status"] == 0 and st[1]["status"] == 0, st
        r0 = np.linalg.norm(oracle.spmv_sub(A, b0, x[:n]))
        assert r0 <= 1e-10 * np.linalg.norm(b0) * (1 + 1e-9)
        if name == "zero":
            assert st[1]["iterations"] == 0
            assert np.all(x[n:] == 0.0)
        else:
            r1 = np.linalg.norm(oracle.spmv_sub(A, second, x[n:]))
            assert r1 <= 1e-10 * np.linalg.norm(second) * (1 + 1e-9)


def test_cg_reference_semantics(gpu, oracle):
    """conjugateGradient: absolute eps, host arrays, x written on early exit,
    max iterations reported with the last iterate (cpp/gpu_simulation_device.cpp:385-515)."""
    from navier_stokes_fem_amd._native import NsfemError, MAX_ITERATIONS_REACHED

    m, A = _cavity_system(oracle, 12, "pressure_stiffness")
    rs, col, val = A
    n = len(rs) - 1
    gpu.upload_csr(12, rs, col, val)
    rng = np.random.default_rng(29)
    b = rng.standard_normal(n)
    x = gpu.conjugate_gradient(12, b, np.zeros(n), -1, 1e-8)
    assert np.linalg.norm(oracle.spmv_sub(A, b, x)) < 1e-8 * (1 + 1e-6)
    # x0 already converged: x must be written (the reference returns early without it)
    x2 = gpu.conjugate_gradient(12, b, x, -1, 1e-3)
    assert np.array_equal(x2, x)
    with pytest.raises(NsfemError) as ei:
        gpu.conjugate_gradient(12, b, np.zeros(n), 3, 1e-12)
    assert ei.value.code == MAX_ITERATIONS_REACHED
    xo, _, _ = oracle.cg(A, b, np.zeros(n), max_iter=3, rtol=0.0, atol=1e-12)
    assert np.allclose(ei.value.x, xo, rtol=1e-10, atol=1e-12)


def test_cg_zero_rhs_and_breakdown(gpu):
    from navier_stokes_fem_amd._native import NsfemError, BREAKDOWN

    n = 50
    rs = np.arange(n + 1, dtype=np.int32)
    col = np.arange(n, dtype=np.int32)
    gpu.upload_csr(13, rs, col, np.full(n, 2.0))
    bx = _buf(gpu, np.zeros(n))
    st = gpu.cg_solve(13, _buf(gpu, np.zeros(n)), bx)[0]
    assert st["iterations"] == 0 and np.all(bx.download() == 0)
    gpu.upload_csr(14, rs, col, -np.ones(n))  # negative definite
    with pytest.raises(NsfemError) as ei:
        gpu.cg_solve(14, _buf(gpu, np.ones(n)), _buf(gpu, np.zeros(n)))
    assert ei.value.code == BREAKDOWN


def test_cg_large_grid(gpu, oracle):
    """A system big enough to run on many persistent blocks (grid barrier)."""
    from navier_stokes_fem_amd import device as D
    from navier_stokes_fem_amd import mesh

    m = mesh.cavity(192)
    a = D.assemble(m, 0.01, 0.01)
    rs, col, val, _ = a["pressure_stiffness"]
    n = len(rs) - 1
    gpu.upload_csr(15, rs, col, val)
    rng = np.random.default_rng(31)
    b = rng.standard_normal(n)
    for precision in ("fp64", "fp32_refined"):
        bx = _buf(gpu, np.zeros(n))
        st = gpu.cg_solve(15, _buf(gpu, b), bx, rtol=1e-8, precision=precision)[0]
        x = bx.download()
        r = np.linalg.norm(oracle.spmv_sub((rs, col, val), b, x))
        assert st["status"] == 0 and r <= 1e-8 * np.linalg.norm(b) * (1 + 1e-9), (precision, st)


def test_cg_two_columns_narrow_equals_packed(gpu, monkeypatch):
    """The narrow-gather two-column iteration stores p_k once where the
    packed one recomputes it from each neighbour's {r, p, q} by the same
    FP32 roundings, with the same products and sums: solutions, iteration
    and refinement counts are bit-identical, on a system of many persistent
    blocks."""
    from navier_stokes_fem_amd import device as D
    from navier_stokes_fem_amd import mesh

    m = mesh.cavity(96)
    a = D.assemble(m, 0.01, 0.01)
    rs, col, val, _ = a["velocity_mass"]
    n = len(rs) - 1
    gpu.upload_csr(15, rs, col, val)
    rng = np.random.default_rng(43)
    b = rng.standard_normal(2 * n)
    out = {}
    for form in ("packed", "narrow"):
        print(f"two columns {form}", flush=True)
        monkeypatch.setenv("NSFEM_CG_PAIRED", form)
        bx = _buf(gpu, np.zeros(2 * n))
        st = gpu.cg_solve(15, _buf(gpu, b), bx, columns=2, rtol=1e-10, precision="fp32_refined")
        out[form] = (bx.download(), [(c["iterations"], c["refinements"], c["status"]) for c in st])
    np.testing.assert_array_equal(out["packed"][0], out["narrow"][0])
    assert out["packed"][1] == out["narrow"][1]
    assert all(c[2] == 0 for c in out["narrow"][1])


@pytest.mark.parametrize("env", [{}, {"NSFEM_CG_PAIRED": "packed"}, {"NSFEM_CG_PAIRED": "twophase"}],
                         ids=["narrow", "packed", "twophase"])
def test_cg_two_columns_long_row(gpu, oracle, monkeypatch, env):
    """Two columns on an SPD arrow matrix whose first row and column are
    dense (3,000 nonzeros, longer than a wave tile, summed wave-wide) around
    a tridiagonal block: every two-column form reaches the tolerance, FP64
    and FP32 + refinement."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    n = 3000
    a = 0.5 / np.sqrt(n)  # eigenvalues in [0.75, 3.0]
    rows, cols, vals = [0] * n, list(range(n)), [2.0] + [a] * (n - 1)
    for j in range(1, n):
        rows += [j, j]
        cols += [0, j]
        vals += [a, 2.0]
        if j > 1:
            rows.append(j)
            cols.append(j - 1)
            vals.append(-0.5)
        if j < n - 1:
            rows.append(j)
            cols.append(j + 1)
            vals.append(-0.5)
    order = np.lexsort((np.array(cols), np.array(rows)))
    r_, c_, v_ = np.array(rows)[order], np.array(cols, np.int32)[order], np.array(vals)[order]
    rs = np.zeros(n + 1, np.int32)
    np.add.at(rs, r_ + 1, 1)
    rs = np.cumsum(rs).astype(np.int32)
    A = (rs, c_, v_)
    assert rs[1] - rs[0] == n  # the dense row
    gpu.upload_csr(14, rs, c_, v_)
    rng = np.random.default_rng(47)
    b = rng.standard_normal(2 * n)
    for precision in ("fp64", "fp32_refined"):
        print(f"arrow matrix, two columns, {precision}, {env}", flush=True)
        bx = _buf(gpu, np.zeros(2 * n))
        st = gpu.cg_solve(14, _buf(gpu, b), bx, columns=2, rtol=1e-10, precision=precision)
        x = bx.download()
        for c in range(2):
            bc = b[c * n:(c + 1) * n]
            assert st[c]["status"] == 0, st
            r = np.linalg.norm(oracle.spmv_sub(A, bc, x[c * n:(c + 1) * n]))
            assert r <= 1e-10 * np.linalg.norm(bc) * (1 + 1e-9)
