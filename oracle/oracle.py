"""ctypes front-end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.  It also exposes
the reference's own KD-tree traversal compiled into oracle/_ref (when that
library was built in a container holding the reference checkout).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_kdtree.so")

_P = np.ctypeslib.ndpointer
_f64 = _P(dtype=np.float64, flags="C_CONTIGUOUS")
_f32 = _P(dtype=np.float32, flags="C_CONTIGUOUS")
_i32 = _P(dtype=np.int32, flags="C_CONTIGUOUS")
_u32 = _P(dtype=np.uint32, flags="C_CONTIGUOUS")

_lib = None
_ref = None


def build(force: bool = False) -> None:
    """Compile the oracle (and oracle/_ref when /root/reference exists)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "-s"], check=True)


class CgResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("rhs_norm", C.c_double),
                ("initial_residual", C.c_double), ("final_residual", C.c_double),
                ("recursive_residual", C.c_double)]


class KdTree(C.Structure):
    _fields_ = [("node_count", C.c_int32), ("nodes", C.POINTER(C.c_uint32)),
                ("leaf_index_count", C.c_int64), ("leaf_indices", C.POINTER(C.c_int32)),
                ("bbox", C.c_float * 4), ("max_depth", C.c_int32)]


class Csr(C.Structure):
    _fields_ = [("rows", C.c_int), ("cols", C.c_int), ("nnz", C.c_int64),
                ("row_start", C.POINTER(C.c_int32)), ("col", C.POINTER(C.c_int32)),
                ("val", C.POINTER(C.c_double))]


class Matrices(C.Structure):
    _fields_ = [(n, Csr) for n in ("velocity_mass", "diffusion", "velocity_weights",
                                   "pressure_stiffness", "pressure_weights",
                                   "velocity_divergence", "pressure_divergence")]


class StepStats(C.Structure):
    _fields_ = [("cg_iterations", C.c_int * 5), ("cg_residual", C.c_double * 5),
                ("status", C.c_int)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_set_threads.argtypes = [C.c_int]
        L.or_get_threads.restype = C.c_int
        L.or_spmv.argtypes = [C.c_int, _i32, _i32, _f64, _f64, _f64]
        L.or_spmv_sub.argtypes = [C.c_int, _i32, _i32, _f64, _f64, _f64, _f64]
        L.or_dot.argtypes = [C.c_int64, _f64, _f64]
        L.or_dot.restype = C.c_double
        L.or_axpy.argtypes = [C.c_int64, C.c_double, _f64, _f64]
        L.or_axpby.argtypes = [C.c_int64, C.c_double, C.c_double, _f64, _f64, _f64]
        L.or_triplet_to_csr.argtypes = [C.c_int, C.c_int64, _i32, _i32, _f64, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.POINTER(C.c_int64)]
        L.or_cg.argtypes = [C.c_int, _i32, _i32, _f64, _f64, _f64, _f64, C.c_int, C.c_double,
                            C.c_double, C.POINTER(CgResult)]
        L.or_grid_bbox.argtypes = [C.c_int, _f32, _f32]
        L.or_kd_build.argtypes = [C.c_int, _f32, C.c_int, _i32, C.c_int, C.c_int, C.POINTER(KdTree)]
        L.or_kd_free.argtypes = [C.POINTER(KdTree)]
        L.or_point_in_triangle.argtypes = [C.c_float] * 8 + [C.POINTER(C.c_float)] * 2
        L.or_kd_find_many.argtypes = [_u32, _i32, _f32, _i32, C.c_int, _f32, _i32, _f32, _f32,
                                      _i32, C.c_int]
        L.or_advect_f32.argtypes = [_u32, _i32, _f32, _i32, C.c_int, _f32, _f32, C.c_float,
                                    _f32, _f32, C.c_int]
        L.or_advect_f64.argtypes = [_u32, _i32, _f32, _i32, C.c_int, _f64, _f64, C.c_double,
                                    _f64, _f64, C.c_int]
        L.or_assemble.argtypes = [C.c_int, C.c_int, _f32, C.c_int, _i32, C.c_int, _i32, C.c_int,
                                  _i32, C.c_double, C.c_double, C.POINTER(Matrices)]
        L.or_matrices_free.argtypes = [C.POINTER(Matrices)]
        L.or_step.argtypes = [C.POINTER(Matrices), _u32, _i32, _f32, _i32, C.c_int, C.c_int,
                              C.c_int, _i32, _f64, _f64, C.c_int, _i32, _f64, C.c_double,
                              C.c_double, C.c_int, _f64, _f64, _f64, C.POINTER(StepStats)]
        L.or_impose_velocity_bc.argtypes = [C.c_int, _i32, _f64, _f64, _f64, _f64]
        L.or_set_step_fault.argtypes = [C.c_int]
        L.or_get_step_fault.restype = C.c_int
        _lib = L
    return _lib


def ref_lib():
    """The reference's own KD traversal (oracle/_ref), or None if absent."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = C.CDLL(REF_PATH)
        R.ref_point_in_triangle.argtypes = [C.c_float] * 8 + [C.POINTER(C.c_float)] * 2
        R.ref_point_in_triangle.restype = C.c_int
        R.ref_kd_find_many.argtypes = [_u32, _i32, _f32, _i32, C.c_int, _f32, C.c_int, _f32,
                                       _i32, _f32, _f32, _i32]
        _ref = R
    return _ref


def set_threads(n: int) -> None:
    lib().or_set_threads(int(n))


def get_threads() -> int:
    return int(lib().or_get_threads())


# ---------------------------------------------------------------------------
def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def spmv(A, x):
    rs, col, val = A
    y = np.empty(len(rs) - 1, dtype=np.float64)
    lib().or_spmv(len(rs) - 1, _c(rs, np.int32), _c(col, np.int32), _c(val, np.float64),
                  _c(x, np.float64), y)
    return y


def spmv_sub(A, lhs, x):
    rs, col, val = A
    y = np.empty(len(rs) - 1, dtype=np.float64)
    lib().or_spmv_sub(len(rs) - 1, _c(rs, np.int32), _c(col, np.int32), _c(val, np.float64),
                      _c(lhs, np.float64), _c(x, np.float64), y)
    return y


def dot(a, b):
    return float(lib().or_dot(len(a), _c(a, np.float64), _c(b, np.float64)))


def axpy(a, x, y):
    y = _c(y, np.float64).copy()
    lib().or_axpy(len(y), float(a), _c(x, np.float64), y)
    return y


def axpby(a, b, x, y):
    out = np.empty(len(x), dtype=np.float64)
    lib().or_axpby(len(x), float(a), float(b), _c(x, np.float64), _c(y, np.float64), out)
    return out


def triplet_to_csr(rows, ti, tj, tv):
    ti, tj, tv = _c(ti, np.int32), _c(tj, np.int32), _c(tv, np.float64)
    nnz = C.c_int64(0)
    rs = np.empty(rows + 1, dtype=np.int32)
    rc = lib().or_triplet_to_csr(rows, len(ti), ti, tj, tv, rs.ctypes.data, None, None,
                                 C.byref(nnz))
    if rc:
        raise ValueError(f"or_triplet_to_csr failed: {rc}")
    col = np.empty(max(nnz.value, 1), dtype=np.int32)
    val = np.empty(max(nnz.value, 1), dtype=np.float64)
    lib().or_triplet_to_csr(rows, len(ti), ti, tj, tv, rs.ctypes.data, col.ctypes.data,
                            val.ctypes.data, C.byref(nnz))
    return rs, col[:nnz.value], val[:nnz.value]


def cg(A, b, x0, max_iter=-1, rtol=1e-8, atol=0.0):
    rs, col, val = A
    x = np.empty(len(b), dtype=np.float64)
    res = CgResult()
    st = lib().or_cg(len(rs) - 1, _c(rs, np.int32), _c(col, np.int32), _c(val, np.float64),
                     _c(b, np.float64), _c(x0, np.float64), x, int(max_iter), float(rtol),
                     float(atol), C.byref(res))
    return x, st, res


class Tree:
    """A KD tree in the reference's node layout (numpy copies)."""

    def __init__(self, nodes, leaf_indices, bbox, max_depth):
        self.nodes = nodes
        self.leaf_indices = leaf_indices
        self.bbox = bbox
        self.max_depth = max_depth


def grid_bbox(xy):
    b = np.empty(4, dtype=np.float32)
    lib().or_grid_bbox(len(xy) // 2, _c(xy, np.float32), b)
    return b


def kd_build(xy, elems, max_depth=-1, min_leaf_size=16) -> Tree:
    t = KdTree()
    xy, elems = _c(xy, np.float32), _c(elems, np.int32)
    rc = lib().or_kd_build(len(xy) // 2, xy, len(elems) // 6, elems, int(max_depth),
                           int(min_leaf_size), C.byref(t))
    if rc:
        raise RuntimeError("or_kd_build failed")
    nodes = np.ctypeslib.as_array(t.nodes, shape=(2 * t.node_count,)).copy()
    leaf = (np.ctypeslib.as_array(t.leaf_indices, shape=(t.leaf_index_count,)).copy()
            if t.leaf_index_count else np.zeros(0, np.int32))
    out = Tree(nodes, leaf, np.array(t.bbox[:], dtype=np.float32), t.max_depth)
    lib().or_kd_free(C.byref(t))
    return out


def point_in_triangle(p, a, b, c):
    xi, eta = C.c_float(0), C.c_float(0)
    ok = lib().or_point_in_triangle(*map(float, (p[0], p[1], a[0], a[1], b[0], b[1], c[0], c[1])),
                                    C.byref(xi), C.byref(eta))
    return bool(ok), xi.value, eta.value


def kd_find(tree: Tree, xy, elems, pts, float_offset=False):
    pts = _c(pts, np.float32).reshape(-1)
    n = len(pts) // 2
    e = np.empty(n, np.int32)
    xi = np.empty(n, np.float32)
    eta = np.empty(n, np.float32)
    nn = np.empty(n, np.int32)
    lib().or_kd_find_many(_c(tree.nodes, np.uint32), _c(tree.leaf_indices, np.int32),
                          _c(xy, np.float32), _c(elems, np.int32), n, pts, e, xi, eta, nn,
                          int(bool(float_offset)))
    return e, xi, eta, nn


def ref_kd_find(tree: Tree, xy, elems, pts):
    R = ref_lib()
    if R is None:
        raise RuntimeError("oracle/_ref/libref_kdtree.so not built")
    pts = _c(pts, np.float32).reshape(-1)
    n = len(pts) // 2
    e = np.empty(n, np.int32)
    xi = np.empty(n, np.float32)
    eta = np.empty(n, np.float32)
    nn = np.empty(n, np.int32)
    R.ref_kd_find_many(_c(tree.nodes, np.uint32), _c(tree.leaf_indices, np.int32),
                       _c(xy, np.float32), _c(elems, np.int32), len(xy) // 2,
                       _c(tree.bbox, np.float32), n, pts, e, xi, eta, nn)
    return e, xi, eta, nn


def advect_f32(tree: Tree, xy, elems, u, v, dt, float_offset=False):
    n = len(u)
    uo = np.empty(n, np.float32)
    vo = np.empty(n, np.float32)
    lib().or_advect_f32(_c(tree.nodes, np.uint32), _c(tree.leaf_indices, np.int32),
                        _c(xy, np.float32), _c(elems, np.int32), n, _c(u, np.float32),
                        _c(v, np.float32), float(dt), uo, vo, int(bool(float_offset)))
    return uo, vo


def advect_f64(tree: Tree, xy, elems, u, v, dt, float_offset=False):
    n = len(u)
    uo = np.empty(n, np.float64)
    vo = np.empty(n, np.float64)
    lib().or_advect_f64(_c(tree.nodes, np.uint32), _c(tree.leaf_indices, np.int32),
                        _c(xy, np.float32), _c(elems, np.int32), n, _c(u, np.float64),
                        _c(v, np.float64), float(dt), uo, vo, int(bool(float_offset)))
    return uo, vo


def _csr_np(c: Csr):
    rs = np.ctypeslib.as_array(c.row_start, shape=(c.rows + 1,)).copy()
    if c.nnz:
        col = np.ctypeslib.as_array(c.col, shape=(c.nnz,)).copy()
        val = np.ctypeslib.as_array(c.val, shape=(c.nnz,)).copy()
    else:
        col, val = np.zeros(0, np.int32), np.zeros(0, np.float64)
    return rs, col, val


class Assembly:
    """Oracle-assembled matrices.  Keeps the C struct alive for or_step."""

    NAMES = ("velocity_mass", "diffusion", "velocity_weights", "pressure_stiffness",
             "pressure_weights", "velocity_divergence", "pressure_divergence")

    def __init__(self, mesh, viscosity, dt):
        self._m = Matrices()
        rc = lib().or_assemble(mesh.node_count, mesh.pressure_count, mesh.xy(),
                               mesh.element_count, mesh.elems(), len(mesh.vel_bc_nodes),
                               _c(mesh.vel_bc_nodes, np.int32), len(mesh.p_bc_nodes),
                               _c(mesh.p_bc_nodes, np.int32), float(viscosity), float(dt),
                               C.byref(self._m))
        if rc:
            raise RuntimeError("or_assemble failed")
        self.csr = {n: _csr_np(getattr(self._m, n)) for n in self.NAMES}

    def __del__(self):
        try:
            lib().or_matrices_free(C.byref(self._m))
        except Exception:
            pass


def step(asm: Assembly, tree: Tree, mesh, dt, u, v, p, rtol=1e-8, max_iter=-1):
    """One oracle update step on copies of (u, v, p); returns (u, v, p, stats)."""
    u = _c(u, np.float64).copy()
    v = _c(v, np.float64).copy()
    p = _c(p, np.float64).copy()
    st = StepStats()
    lib().or_step(C.byref(asm._m), _c(tree.nodes, np.uint32), _c(tree.leaf_indices, np.int32),
                  mesh.xy(), mesh.elems(), mesh.node_count, mesh.pressure_count,
                  len(mesh.vel_bc_nodes), _c(mesh.vel_bc_nodes, np.int32),
                  _c(mesh.vel_bc_u, np.float64), _c(mesh.vel_bc_v, np.float64),
                  len(mesh.p_bc_nodes), _c(mesh.p_bc_nodes, np.int32),
                  _c(mesh.p_bc_values, np.float64), float(dt), float(rtol), int(max_iter),
                  u, v, p, C.byref(st))
    stats = {"iterations": list(st.cg_iterations), "residual": list(st.cg_residual),
             "status": st.status}
    return u, v, p, stats


# Planted mistakes of or_step (OR_FAULT_* of nsfem_oracle.h), for the tests
# that show the composition checks catch them.
STEP_FAULTS = ("div_sign", "grad_sign", "vel_bc_sign", "p_bc_sign", "div_scale",
               "grad_scale", "no_update_add", "skip_advection", "swap_components")


class step_fault:
    """Context manager: or_step commits the named mistake inside the block."""

    def __init__(self, name):
        self.code = 0 if name in (None, "none") else STEP_FAULTS.index(name) + 1

    def __enter__(self):
        lib().or_set_step_fault(self.code)
        return self

    def __exit__(self, *exc):
        lib().or_set_step_fault(0)
        return False


def initial_fields(mesh):
    """Zero fields with the velocity Dirichlet values imposed
    (include/assembly.h:1503-1507)."""
    V, P = mesh.node_count, mesh.pressure_count
    u = np.zeros(V)
    v = np.zeros(V)
    lib().or_impose_velocity_bc(len(mesh.vel_bc_nodes), _c(mesh.vel_bc_nodes, np.int32),
                                _c(mesh.vel_bc_u, np.float64), _c(mesh.vel_bc_v, np.float64),
                                u, v)
    return u, v, np.zeros(P)
