"""CPU-side checks of the product library and its host code (no GPU calls):

* libnsfem_hip.so loads and exports every function include/nsfem_gpu.h
  declares, with the status codes the header defines;
* the product's host KD-tree build (KDTreeBuilder::build,
  cpp/kd_tree_builder.cpp:81-148) and FEM assembly (include/assembly.h:
  1302-1475) are bit-identical to the oracle's;
* the mesh generators produce the layouts the kernels assume;
* bench.py refuses more than one GPU before touching a device.
"""
import ctypes as C
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "nsfem_gpu.h")


def _header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nsfem_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    from navier_stokes_fem_amd import _native as N

    lib = C.CDLL(N.LIB_PATH)
    names = _header_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes layer binds exactly the declared set
    assert sorted(N.EXPORTED) == names


def test_status_codes_match_header():
    from navier_stokes_fem_amd import _native as N

    text = open(HEADER).read()
    codes = dict((k, int(v)) for k, v in re.findall(r"#define NSFEM_([A-Z_0-9]+)\s+\(?(-?\d+)\)?", text))
    for name in ("OK", "ERROR", "MAX_ITERATIONS_REACHED", "INVALID_ARGUMENT", "HIP_ERROR",
                 "OUT_OF_MEMORY", "NOT_READY", "BREAKDOWN", "BARRIER_TIMEOUT", "STAGNATION"):
        assert getattr(N, name) == codes[name], name
    assert N.PRECISION_FP64 == codes["PRECISION_FP64"]
    assert N.PRECISION_FP32_REFINED == codes["PRECISION_FP32_REFINED"]


def test_host_argument_errors_are_reported():
    """Host-side validation: bad arguments return INVALID_ARGUMENT with a
    message, before any device work."""
    from navier_stokes_fem_amd import _native as N

    L = N.lib()
    h = C.c_void_p()
    rc = L.nsfem_kdtree_build(-1, None, 0, None, -1, 16, C.byref(h))
    assert rc == N.INVALID_ARGUMENT
    assert L.nsfem_last_error()


def _meshes(mesh_small):
    from navier_stokes_fem_amd import mesh

    return [mesh.cavity(12), mesh.channel(nx=40, ny=10, seed=4), mesh_small]


def test_host_kdtree_matches_oracle(oracle, mesh_small):
    from navier_stokes_fem_amd import device as D

    for m in _meshes(mesh_small):
        a = D.build_kdtree(m)
        b = oracle.kd_build(m.xy(), m.elems())
        np.testing.assert_array_equal(a.nodes, b.nodes)
        np.testing.assert_array_equal(a.leaf_indices, b.leaf_indices)
        np.testing.assert_array_equal(a.bbox, b.bbox)
        assert a.max_depth == b.max_depth


def test_host_assembly_matches_oracle(oracle, mesh_small):
    from navier_stokes_fem_amd import device as D

    for m in _meshes(mesh_small):
        a = D.assemble(m, 0.01, 0.005, threads=3)
        o = oracle.Assembly(m, 0.01, 0.005)
        for name in ("velocity_mass", "diffusion", "pressure_stiffness",
                     "velocity_divergence", "pressure_divergence"):
            rs, col, val, _ = a[name]
            ors, ocol, oval = o.csr[name]
            np.testing.assert_array_equal(rs, ors, err_msg=name)
            np.testing.assert_array_equal(col, ocol, err_msg=name)
            np.testing.assert_array_equal(val, oval, err_msg=name)
        # Dirichlet corrections: c = W^T g with g summed per listing
        # (include/assembly.h:503-524; W = velocity/pressure weights)
        V, P = m.node_count, m.pressure_count
        gu, gv, gp = np.zeros(V), np.zeros(V), np.zeros(P)
        np.add.at(gu, m.vel_bc_nodes, m.vel_bc_u)
        np.add.at(gv, m.vel_bc_nodes, m.vel_bc_v)
        np.add.at(gp, m.p_bc_nodes, m.p_bc_values)
        import scipy.sparse as sp

        def csr(name, n):
            rs, col, val = o.csr[name]
            return sp.csr_matrix((val, col, rs), shape=(n, n))

        W, Wp = csr("velocity_weights", V), csr("pressure_weights", P)
        np.testing.assert_allclose(a["c_u"], W.T @ gu, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(a["c_v"], W.T @ gv, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(a["c_p"], Wp.T @ gp, rtol=1e-12, atol=1e-14)


def test_host_assembly_thread_count_invariant(mesh_small):
    from navier_stokes_fem_amd import device as D, mesh

    m = mesh.channel(nx=30, ny=8, seed=1)
    a1 = D.assemble(m, 0.01, 0.01, threads=1)
    a8 = D.assemble(m, 0.01, 0.01, threads=8)
    for k in a1:
        for u, v in zip(a1[k] if isinstance(a1[k], tuple) else [a1[k]],
                        a8[k] if isinstance(a8[k], tuple) else [a8[k]]):
            np.testing.assert_array_equal(np.asarray(u), np.asarray(v))


# ------------------------------------------------------------------ meshes

def _signed_areas(m):
    xy = m.nodes.astype(np.float64)
    e = m.elements
    a, b, c = xy[e[:, 0]], xy[e[:, 1]], xy[e[:, 2]]
    return 0.5 * ((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0]))


def test_cavity_layout():
    from navier_stokes_fem_amd import mesh

    n = 8
    m = mesh.cavity(n, lid_speed=2.0)
    m.check()
    assert m.node_count == (2 * n + 1) ** 2 and m.pressure_count == (n + 1) ** 2
    # counter-clockwise triangles only (the point test is CCW-only)
    assert np.all(_signed_areas(m) > 0)
    np.testing.assert_allclose(_signed_areas(m).sum(), 1.0, rtol=1e-6)
    # midpoints sit at edge midpoints (FP32)
    xy = m.nodes.astype(np.float64)
    e = m.elements
    for k, (a, b) in zip((3, 4, 5), ((1, 2), (0, 2), (0, 1))):
        np.testing.assert_allclose(xy[e[:, k]], 0.5 * (xy[e[:, a]] + xy[e[:, b]]), atol=1e-7)
    # lid: u = lid speed on y = 1 except the corners; no-slip elsewhere
    bxy = xy[m.vel_bc_nodes]
    lid = (bxy[:, 1] == 1.0) & (bxy[:, 0] > 0) & (bxy[:, 0] < 1)
    np.testing.assert_array_equal(m.vel_bc_u[lid], 2.0)
    np.testing.assert_array_equal(m.vel_bc_u[~lid], 0.0)
    np.testing.assert_array_equal(m.vel_bc_v, 0.0)
    on_boundary = (xy[:, 0] == 0) | (xy[:, 0] == 1) | (xy[:, 1] == 0) | (xy[:, 1] == 1)
    assert set(m.vel_bc_nodes.tolist()) == set(np.flatnonzero(on_boundary).tolist())
    assert m.p_bc_nodes.tolist() == [0] and m.p_bc_values.tolist() == [0.0]


def test_channel_layout():
    from navier_stokes_fem_amd import mesh

    nx, ny, L, H = 64, 16, 2.2, 0.41
    m = mesh.channel(nx=nx, ny=ny, seed=2)
    m.check()
    assert np.all(_signed_areas(m) > 0)
    xy = m.nodes.astype(np.float64)
    # seeded: the same seed gives the same mesh, another seed another one
    np.testing.assert_array_equal(m.nodes, mesh.channel(nx=nx, ny=ny, seed=2).nodes)
    assert not np.array_equal(m.nodes, mesh.channel(nx=nx, ny=ny, seed=3).nodes)
    # vertices stay within +-0.2 h of their lattice position, inside the box
    P = m.pressure_count
    assert xy[:P, 0].min() == 0.0 and xy[:P, 0].max() == pytest.approx(L)
    assert xy[:P, 1].min() == 0.0 and xy[:P, 1].max() == pytest.approx(H)
    hx, hy = L / nx, H / ny
    fx = xy[:P, 0] / hx - np.round(xy[:P, 0] / hx)
    fy = xy[:P, 1] / hy - np.round(xy[:P, 1] / hy)
    assert np.abs(fx).max() <= 0.2 + 1e-5 and np.abs(fy).max() <= 0.2 + 1e-5
    # no triangle centroid inside the disk
    c = xy[m.elements[:, :3]].mean(1)
    assert np.all(((c[:, 0] - 0.2) ** 2 + (c[:, 1] - 0.2) ** 2) >= 0.05 ** 2)
    assert m.element_count < 2 * nx * ny
    # inflow profile on x = 0, outlet pressure on x = L
    b = xy[m.vel_bc_nodes]
    inflow = b[:, 0] == 0.0
    np.testing.assert_allclose(m.vel_bc_u[inflow], 4 * 1.5 * b[inflow, 1] * (H - b[inflow, 1]) / H ** 2,
                               rtol=1e-12, atol=1e-12)
    np.testing.assert_array_equal(m.vel_bc_u[~inflow], 0.0)
    assert np.all(xy[m.p_bc_nodes, 0] == pytest.approx(L))
    np.testing.assert_array_equal(m.p_bc_values, 0.0)


def test_expression_parser():
    from navier_stokes_fem_amd.mesh import evaluate_expression as ev

    x, y = np.array([0.0, 1.0, 2.0]), np.array([0.5, 0.25, -1.0])
    np.testing.assert_allclose(ev("4*1.5*y*(0.41-y)/(0.41*0.41)", x, y),
                               6.0 * y * (0.41 - y) / 0.41 ** 2)
    np.testing.assert_allclose(ev("1 - 2 - 3", x, y), -4.0)      # left associative
    np.testing.assert_allclose(ev("8 / 4 / 2", x, y), 1.0)
    np.testing.assert_allclose(ev("-x^2", x, y), -(x ** 2))       # ^ binds tighter
    np.testing.assert_allclose(ev("2*x + 3*y", x, y), 2 * x + 3 * y)
    np.testing.assert_allclose(ev("sin(x) + cos(y)", x, y), np.sin(x) + np.cos(y))
    np.testing.assert_allclose(ev("0", x, y), 0.0)


def test_load_json_roundtrip(tmp_path, mesh_small):
    """The reference's JSON schema (cpp/grid.cpp:19-72) on a synthetic file."""
    from navier_stokes_fem_amd import mesh

    m = mesh.cavity(2)
    xy = m.nodes.astype(np.float64)
    top = np.flatnonzero(xy[:, 1] == 1.0).tolist()
    d = {"elementsCount": m.element_count, "velocityNodesCount": m.node_count,
         "pressureNodesCount": m.pressure_count, "elementSize": 6,
         "nodes": m.nodes.astype(np.float64).tolist(), "elements": m.elements.reshape(-1).tolist(),
         "uDirichlet": [{"nodes": top, "u": "2*x", "v": "0"}],
         "pDirichlet": [{"nodes": [0], "p": "1+1"}]}
    p = tmp_path / "m.json"
    p.write_text(json.dumps(d))
    r = mesh.load_json(str(p))
    np.testing.assert_array_equal(r.nodes, m.nodes)
    np.testing.assert_array_equal(r.elements, m.elements)
    np.testing.assert_allclose(r.vel_bc_u, 2 * xy[top, 0])
    np.testing.assert_array_equal(r.p_bc_values, [2.0])


def test_bench_refuses_several_gpus():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 2
    line = json.loads(out.stdout.strip().splitlines()[-1])
    assert "one GPU" in line.get("error", "") or "single GPU" in line.get("error", "")


def test_cpp_device_header_compiles(tmp_path):
    """include/nsfem_gpu_device.hpp is self-contained C++17 over the C ABI."""
    src = tmp_path / "t.cpp"
    src.write_text('#include "nsfem_gpu_device.hpp"\n'
                   "int main() { nsfem::GPUSimulationDevice d; (void)d; return 0; }\n")
    out = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                          "-I" + os.path.join(ROOT, "include"), str(src)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
