"""Python mirror of the reference's GPU simulation device
(include/gpu_simulation_device.h:13-206), backed by libnsfem_hip.so.

Method names follow the reference (``uploadMatrix`` -> :meth:`upload_matrix`,
``spRMult`` -> :meth:`sp_r_mult`, ...); failures raise
:class:`~navier_stokes_fem_amd._native.NsfemError` carrying the status the
reference's ``EC::ErrorCode`` would hold.  Arrays are numpy; device vectors are
:class:`DeviceBuffer` objects (the reference's ``GPU::GPUBuffer``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

__all__ = ["GPUSimulationDevice", "DeviceBuffer", "KdTree", "build_kdtree", "assemble",
           "device_count"]

PRESSURE_STIFFNESS = 0
VELOCITY_MASS = 1
DIFFUSION = 2
VELOCITY_DIVERGENCE = 3
PRESSURE_DIVERGENCE = 4

_PREC = {"fp64": N.PRECISION_FP64, "fp32_refined": N.PRECISION_FP32_REFINED,
         N.PRECISION_FP64: N.PRECISION_FP64, N.PRECISION_FP32_REFINED: N.PRECISION_FP32_REFINED}


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def device_count() -> int:
    n = C.c_int(0)
    N.check(N.lib().nsfem_get_device_count(C.byref(n)))
    return n.value


class DeviceBuffer:
    """Device memory owned by a device handle (GPU::GPUBuffer)."""

    def __init__(self, device: "GPUSimulationDevice", nbytes: int):
        self.device = device
        h = C.c_void_p()
        N.check(N.lib().nsfem_buffer_create(device.handle, int(nbytes), C.byref(h)))
        self.handle = h
        self.nbytes = int(nbytes)

    @classmethod
    def from_array(cls, device, a):
        a = np.ascontiguousarray(a)
        b = cls(device, a.nbytes)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a)
        N.check(N.lib().nsfem_buffer_upload(self.handle, N.ptr(a), a.nbytes))

    def download(self, dtype=np.float64, count=None):
        itemsize = np.dtype(dtype).itemsize
        count = self.nbytes // itemsize if count is None else int(count)
        out = np.empty(count, dtype=dtype)
        N.check(N.lib().nsfem_buffer_download(self.handle, N.ptr(out), out.nbytes))
        return out

    def close(self):
        if getattr(self, "handle", None):
            N.lib().nsfem_buffer_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class KdTree:
    """A KD tree in the reference layout (KDTreeCPUOwner): node words (2 per
    KDNode), leaf triangle indices, bbox, max depth."""

    def __init__(self, nodes, leaf_indices, bbox, max_depth):
        self.nodes = nodes
        self.leaf_indices = leaf_indices
        self.bbox = bbox
        self.max_depth = max_depth

    @property
    def node_count(self):
        return len(self.nodes) // 2


def build_kdtree(mesh, max_depth: int = -1, min_leaf_size: int = 16) -> KdTree:
    """KDTreeBuilder::buildCPUOwner (cpp/kd_tree_builder.cpp:17-25) on the host."""
    xy, el = mesh.xy(), mesh.elems()
    L = N.lib()
    h = C.c_void_p()
    N.check(L.nsfem_kdtree_build(mesh.node_count, N.ptr(xy), mesh.element_count, N.ptr(el),
                                 int(max_depth), int(min_leaf_size), C.byref(h)))
    try:
        nc, lc, md = C.c_int32(0), C.c_int64(0), C.c_int32(0)
        bbox = np.zeros(4, np.float32)
        N.check(L.nsfem_kdtree_info(h, C.byref(nc), C.byref(lc), C.byref(md), N.ptr(bbox)))
        nodes = np.empty(2 * nc.value, np.uint32)
        leaf = np.empty(max(lc.value, 1), np.int32)
        N.check(L.nsfem_kdtree_copy(h, N.ptr(nodes), N.ptr(leaf)))
        return KdTree(nodes, leaf[:lc.value], bbox, md.value)
    finally:
        L.nsfem_kdtree_destroy(h)


def assemble(mesh, viscosity: float, dt: float, threads: int = 0, device=None) -> dict:
    """assembleAllMatrices (include/assembly.h:1302-1475), FP64: on the host
    threads, or on the GPU of ``device`` (a :class:`GPUSimulationDevice`;
    bit-identical results).  Returns CSR triples (row_start, col, val) for
    ``velocity_mass``, ``diffusion``, ``pressure_stiffness``,
    ``velocity_divergence``, ``pressure_divergence`` and the Dirichlet
    corrections ``c_u, c_v, c_p``."""
    L = N.lib()
    h = C.c_void_p()
    xy, el = mesh.xy(), mesh.elems()
    vb, gu, gv = _c(mesh.vel_bc_nodes, np.int32), _c(mesh.vel_bc_u, np.float64), _c(mesh.vel_bc_v, np.float64)
    pb, gp = _c(mesh.p_bc_nodes, np.int32), _c(mesh.p_bc_values, np.float64)
    if device is None:
        N.check(L.nsfem_assemble(mesh.node_count, mesh.pressure_count, N.ptr(xy),
                                 mesh.element_count, N.ptr(el), len(vb), N.ptr(vb), N.ptr(gu),
                                 N.ptr(gv), len(pb), N.ptr(pb), N.ptr(gp), float(viscosity),
                                 float(dt), int(threads), C.byref(h)))
    else:
        N.check(L.nsfem_assemble_device(device.handle, mesh.node_count, mesh.pressure_count,
                                        N.ptr(xy), mesh.element_count, N.ptr(el), len(vb),
                                        N.ptr(vb), N.ptr(gu), N.ptr(gv), len(pb), N.ptr(pb),
                                        N.ptr(gp), float(viscosity), float(dt), C.byref(h)))
    try:
        out = {}
        names = ("velocity_mass", "diffusion", "pressure_stiffness", "velocity_divergence",
                 "pressure_divergence")
        for k, name in enumerate(names):
            r, c, nnz = C.c_int(0), C.c_int(0), C.c_int64(0)
            N.check(L.nsfem_assembly_matrix_info(h, k, C.byref(r), C.byref(c), C.byref(nnz)))
            rs = np.empty(r.value + 1, np.int32)
            col = np.empty(max(nnz.value, 1), np.int32)
            val = np.empty(max(nnz.value, 1), np.float64)
            N.check(L.nsfem_assembly_matrix_copy(h, k, N.ptr(rs), N.ptr(col), N.ptr(val)))
            out[name] = (rs, col[:nnz.value], val[:nnz.value], (r.value, c.value))
        cu = np.empty(mesh.node_count)
        cv = np.empty(mesh.node_count)
        cp = np.empty(mesh.pressure_count)
        N.check(L.nsfem_assembly_bc_copy(h, N.ptr(cu), N.ptr(cv), N.ptr(cp)))
        out["c_u"], out["c_v"], out["c_p"] = cu, cv, cp
        return out
    finally:
        L.nsfem_assembly_destroy(h)


class GPUSimulationDevice:
    """GPUSimulation::GPUSimulationDevice (include/gpu_simulation_device.h)."""

    pressureSitffness = PRESSURE_STIFFNESS  # the reference's spelling
    PRESSURE_STIFFNESS = PRESSURE_STIFFNESS
    VELOCITY_MASS = VELOCITY_MASS
    DIFFUSION = DIFFUSION
    VELOCITY_DIVERGENCE = VELOCITY_DIVERGENCE
    PRESSURE_DIVERGENCE = PRESSURE_DIVERGENCE

    def __init__(self, index: int = 0):
        h = C.c_void_p()
        N.check(N.lib().nsfem_device_create(int(index), C.byref(h)))
        self.handle = h
        self.index = index
        self._rows = {}

    # -- lifetime -----------------------------------------------------
    def close(self):
        if getattr(self, "handle", None):
            N.lib().nsfem_device_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def info(self) -> dict:
        name = C.create_string_buffer(64)
        cu, hbm, clk = C.c_int(0), C.c_int64(0), C.c_int(0)
        N.check(N.lib().nsfem_device_info(self.handle, name, 64, C.byref(cu), C.byref(hbm),
                                          C.byref(clk)))
        return {"arch": name.value.decode(), "cu_count": cu.value, "hbm_bytes": hbm.value,
                "clock_khz": clk.value}

    def synchronize(self):
        N.check(N.lib().nsfem_device_synchronize(self.handle))

    def buffer(self, a=None, nbytes=None) -> DeviceBuffer:
        if a is not None:
            return DeviceBuffer.from_array(self, a)
        return DeviceBuffer(self, nbytes)

    # -- matrices -----------------------------------------------------
    def upload_matrix(self, slot, rows, cols, row_idx, col_idx, values):
        """uploadMatrix(SimMatrix, TripletMatrix) (cpp/gpu_simulation_device.cpp:133-140)."""
        ri, ci, v = _c(row_idx, np.int32), _c(col_idx, np.int32), _c(values, np.float64)
        N.check(N.lib().nsfem_upload_matrix_triplet(self.handle, int(slot), int(rows), int(cols),
                                                    len(ri), N.ptr(ri), N.ptr(ci), N.ptr(v)))
        self._rows[slot] = (rows, cols)

    def upload_csr(self, slot, row_start, col, val, shape=None):
        rs, c, v = _c(row_start, np.int32), _c(col, np.int32), _c(val, np.float64)
        rows = len(rs) - 1
        cols = shape[1] if shape is not None else rows
        N.check(N.lib().nsfem_upload_matrix_csr(self.handle, int(slot), rows, int(cols), len(c),
                                                N.ptr(rs), N.ptr(c), N.ptr(v)))
        self._rows[slot] = (rows, cols)

    def matrix_info(self, slot):
        r, c, nnz = C.c_int(0), C.c_int(0), C.c_int64(0)
        N.check(N.lib().nsfem_matrix_info(self.handle, int(slot), C.byref(r), C.byref(c),
                                          C.byref(nnz)))
        return r.value, c.value, nnz.value

    # -- CSR products and BLAS-1 (FP64 device buffers) -----------------
    def sp_r_mult(self, slot, mult: DeviceBuffer, res: DeviceBuffer):
        N.check(N.lib().nsfem_sp_r_mult(self.handle, int(slot), mult.handle, res.handle))

    def time_sp_r_mult(self, slot, mult: DeviceBuffer, res: DeviceBuffer, repeats: int = 100):
        """Mean device seconds of one ``sp_r_mult`` launch over `repeats`
        back-to-back launches (HIP events on the handle's stream)."""
        sec = C.c_double(0.0)
        N.check(N.lib().nsfem_time_sp_r_mult(self.handle, int(slot), mult.handle, res.handle,
                                             int(repeats), C.byref(sec)))
        return sec.value

    def sp_r_mult_sub(self, slot, lhs: DeviceBuffer, mult: DeviceBuffer, res: DeviceBuffer):
        N.check(N.lib().nsfem_sp_r_mult_sub(self.handle, int(slot), lhs.handle, mult.handle,
                                            res.handle))

    def saxpy(self, n, a, x: DeviceBuffer, y: DeviceBuffer):
        N.check(N.lib().nsfem_saxpy(self.handle, int(n), float(a), x.handle, y.handle))

    def saxpby(self, n, a, b, x: DeviceBuffer, y: DeviceBuffer, result: DeviceBuffer):
        N.check(N.lib().nsfem_saxpby(self.handle, int(n), float(a), float(b), x.handle, y.handle,
                                     result.handle))

    def dot_product(self, n, a: DeviceBuffer, b: DeviceBuffer, result: DeviceBuffer):
        N.check(N.lib().nsfem_dot_product(self.handle, int(n), a.handle, b.handle, result.handle))

    # -- conjugate gradient -------------------------------------------
    def conjugate_gradient(self, slot, b, x0, max_iterations=-1, eps=1e-8,
                           use_megakernel=True):
        """conjugateGradient (include/gpu_simulation_device.h:62-76): host
        arrays, absolute ``||r|| < eps``.  Returns x; on failure raises
        NsfemError whose ``x`` attribute holds the last iterate."""
        b, x0 = _c(b, np.float64), _c(x0, np.float64)
        x = np.empty_like(b)
        st = N.lib().nsfem_conjugate_gradient(self.handle, int(slot), N.ptr(b), N.ptr(x0),
                                              N.ptr(x), int(max_iterations), float(eps),
                                              int(bool(use_megakernel)))
        if st != N.OK:
            err = N.NsfemError(st, N.last_error())
            err.x = x
            raise err
        return x

    def cg_solve(self, slot, b: DeviceBuffer, x: DeviceBuffer, columns=1, max_iterations=-1,
                 rtol=1e-8, atol=0.0, precision="fp64", raise_on_error=True):
        """Device-resident CG; x holds x0 and receives the solution.  Returns
        a list of per-column stats dicts."""
        stats = (N.CgStats * 2)()
        st = N.lib().nsfem_cg_solve(self.handle, int(slot), b.handle, x.handle, int(columns),
                                    int(max_iterations), float(rtol), float(atol),
                                    _PREC[precision], stats)
        out = [stats[c].as_dict() for c in range(columns)]
        if st != N.OK and raise_on_error:
            err = N.NsfemError(st, N.last_error())
            err.stats = out
            raise err
        return out

    # -- advection ------------------------------------------------------
    def upload_kd_tree(self, tree: KdTree, mesh):
        """uploadKDTree (cpp/gpu_simulation_device.cpp:72-77)."""
        xy, el = mesh.xy(), mesh.elems()
        nodes, leaf = _c(tree.nodes, np.uint32), _c(tree.leaf_indices, np.int32)
        N.check(N.lib().nsfem_upload_kdtree_arrays(self.handle, len(nodes) // 2, N.ptr(nodes),
                                                   len(leaf), N.ptr(leaf) if len(leaf) else None,
                                                   mesh.node_count, N.ptr(xy),
                                                   mesh.element_count, N.ptr(el)))

    def init_velocity_buffers(self, n):
        N.check(N.lib().nsfem_init_velocity_buffers(self.handle, int(n)))

    def advect(self, u, v, dt):
        """advect (cpp/gpu_simulation_device.cpp:91-131): FP32 host arrays."""
        u, v = _c(u, np.float32), _c(v, np.float32)
        uo, vo = np.empty_like(u), np.empty_like(v)
        N.check(N.lib().nsfem_advect(self.handle, len(u), N.ptr(u), N.ptr(v), float(dt),
                                     N.ptr(uo), N.ptr(vo)))
        return uo, vo

    def advect_f64(self, n, u: DeviceBuffer, v: DeviceBuffer, dt, u_out: DeviceBuffer,
                   v_out: DeviceBuffer):
        N.check(N.lib().nsfem_advect_f64(self.handle, int(n), u.handle, v.handle, float(dt),
                                         u_out.handle, v_out.handle))

    def locate(self, points):
        pts = _c(points, np.float32).reshape(-1)
        n = len(pts) // 2
        e = np.empty(n, np.int32)
        xi = np.empty(n, np.float32)
        eta = np.empty(n, np.float32)
        nn = np.empty(n, np.int32)
        N.check(N.lib().nsfem_locate(self.handle, n, N.ptr(pts), N.ptr(e), N.ptr(xi), N.ptr(eta),
                                     N.ptr(nn)))
        return e, xi, eta, nn
