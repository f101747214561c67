"""ctypes binding of libnsfem_hip.so (the C ABI in include/nsfem_gpu.h).

The library is built in-tree (``navier_stokes_fem_amd/libnsfem_hip.so``) by
``navier_stokes_fem_amd.build``; there is no fallback: every GPU entry point
raises if the library is missing or the call fails.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnsfem_hip.so")

OK = 0
ERROR = -1
MAX_ITERATIONS_REACHED = -2
INVALID_ARGUMENT = -3
HIP_ERROR = -4
OUT_OF_MEMORY = -5
NOT_READY = -6
BREAKDOWN = -7
BARRIER_TIMEOUT = -8
STAGNATION = -9

PRECISION_FP64 = 0
PRECISION_FP32_REFINED = 1
ASSEMBLY_DEVICE = 0
ASSEMBLY_HOST = 1

WARM_PREVIOUS = 0
WARM_ZERO = 1
WARM_LEARNED = 2
WARM_LEARNED_OR_PREVIOUS = 3
WARM_LEARNED_CORRECTION = 4
WARM_EXTRAPOLATE = 5
WARM_PROJECTION = 6
WARM_LEARNED_SEARCH = 7
WARM_LEARNED_PROJECTION = 8
GNN_PARAM_COUNT = 4481


class NsfemError(RuntimeError):
    """A failed library call; ``code`` is the nsfem status (EC::ErrorCode
    status of the reference)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message


class CgStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("refinements", C.c_int32), ("status", C.c_int32),
                ("columns", C.c_int32), ("rhs_norm", C.c_double),
                ("initial_residual", C.c_double), ("final_residual", C.c_double),
                ("tolerance", C.c_double), ("seconds", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class SimOptions(C.Structure):
    _fields_ = [("viscosity", C.c_double), ("dt", C.c_double), ("rtol", C.c_double),
                ("max_iterations", C.c_int32), ("precision", C.c_int32),
                ("pair_velocity", C.c_int32), ("threads", C.c_int32), ("assembly", C.c_int32),
                ("velocity_jacobi", C.c_int32)]


class StepStats(C.Structure):
    _fields_ = [("steps", C.c_int32), ("cg_iterations", C.c_int64 * 5),
                ("max_relative_residual", C.c_double * 5), ("seconds", C.c_double),
                ("advect_seconds", C.c_double), ("cg_seconds", C.c_double * 5),
                ("cg_refinements", C.c_int64 * 5), ("coarse_iterations", C.c_int64),
                ("coarse_seconds", C.c_double), ("cg_launches", C.c_int64 * 5),
                ("cg_launch_iterations", C.c_int64 * 5), ("vcycle_solves", C.c_int64 * 5),
                ("vcycle_iterations", C.c_int64 * 5), ("vcycle_seconds", C.c_double * 5)]

    def as_dict(self):
        return {"steps": self.steps, "cg_iterations": list(self.cg_iterations),
                "max_relative_residual": list(self.max_relative_residual),
                "seconds": self.seconds, "advect_seconds": self.advect_seconds,
                "cg_seconds": list(self.cg_seconds),
                "cg_refinements": list(self.cg_refinements),
                "coarse_iterations": self.coarse_iterations,
                "coarse_seconds": self.coarse_seconds,
                "cg_launches": list(self.cg_launches),
                "cg_launch_iterations": list(self.cg_launch_iterations),
                "vcycle_solves": list(self.vcycle_solves),
                "vcycle_iterations": list(self.vcycle_iterations),
                "vcycle_seconds": list(self.vcycle_seconds)}


class BatchTuning(C.Structure):
    _fields_ = [("rows_per_wave", C.c_int32), ("nnz_per_wave", C.c_int32),
                ("block_scale", C.c_int32), ("velocity_block_scale", C.c_int32),
                ("seconds_per_iteration", C.c_double)]

    def as_dict(self):
        return {"rows_per_wave": self.rows_per_wave, "nnz_per_wave": self.nnz_per_wave,
                "block_scale": self.block_scale,
                "velocity_block_scale": self.velocity_block_scale,
                "seconds_per_iteration": self.seconds_per_iteration}


_lib = None
vp = C.c_void_p
i32 = C.c_int32
i64 = C.c_int64
dbl = C.c_double
flt = C.c_float
PP = C.POINTER(C.c_void_p)

_SIGNATURES = {
    "nsfem_last_error": ([], C.c_char_p),
    "nsfem_get_device_count": ([C.POINTER(C.c_int)], C.c_int),
    "nsfem_device_create": ([C.c_int, PP], C.c_int),
    "nsfem_device_destroy": ([vp], C.c_int),
    "nsfem_device_info": ([vp, C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(i64),
                           C.POINTER(C.c_int)], C.c_int),
    "nsfem_device_synchronize": ([vp], C.c_int),
    "nsfem_upload_matrix_triplet": ([vp, C.c_int, C.c_int, C.c_int, i64, vp, vp, vp], C.c_int),
    "nsfem_upload_matrix_csr": ([vp, C.c_int, C.c_int, C.c_int, i64, vp, vp, vp], C.c_int),
    "nsfem_matrix_info": ([vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(i64)],
                          C.c_int),
    "nsfem_free_matrix": ([vp, C.c_int], C.c_int),
    "nsfem_buffer_create": ([vp, i64, PP], C.c_int),
    "nsfem_buffer_destroy": ([vp], C.c_int),
    "nsfem_buffer_upload": ([vp, vp, i64], C.c_int),
    "nsfem_buffer_download": ([vp, vp, i64], C.c_int),
    "nsfem_buffer_byte_size": ([vp], i64),
    "nsfem_buffer_device_pointer": ([vp], vp),
    "nsfem_sp_r_mult": ([vp, C.c_int, vp, vp], C.c_int),
    "nsfem_sp_r_mult_sub": ([vp, C.c_int, vp, vp, vp], C.c_int),
    "nsfem_time_sp_r_mult": ([vp, C.c_int, vp, vp, C.c_int, C.POINTER(C.c_double)], C.c_int),
    "nsfem_saxpy": ([vp, C.c_int, dbl, vp, vp], C.c_int),
    "nsfem_saxpby": ([vp, C.c_int, dbl, dbl, vp, vp, vp], C.c_int),
    "nsfem_dot_product": ([vp, C.c_int, vp, vp, vp], C.c_int),
    "nsfem_conjugate_gradient": ([vp, C.c_int, vp, vp, vp, C.c_int, dbl, C.c_int], C.c_int),
    "nsfem_cg_solve": ([vp, C.c_int, vp, vp, C.c_int, C.c_int, dbl, dbl, C.c_int,
                        C.POINTER(CgStats)], C.c_int),
    "nsfem_cg_solve_noalloc": ([vp, C.c_int, vp, vp, C.c_int, C.c_int, dbl, dbl, C.c_int,
                                C.POINTER(CgStats)], C.c_int),
    "nsfem_kdtree_build": ([C.c_int, vp, C.c_int, vp, C.c_int, C.c_int, PP], C.c_int),
    "nsfem_kdtree_destroy": ([vp], C.c_int),
    "nsfem_kdtree_info": ([vp, C.POINTER(i32), C.POINTER(i64), C.POINTER(i32), vp], C.c_int),
    "nsfem_kdtree_copy": ([vp, vp, vp], C.c_int),
    "nsfem_upload_kdtree": ([vp, vp, C.c_int, vp, C.c_int, vp], C.c_int),
    "nsfem_upload_kdtree_arrays": ([vp, i32, vp, i64, vp, C.c_int, vp, C.c_int, vp], C.c_int),
    "nsfem_init_velocity_buffers": ([vp, C.c_int], C.c_int),
    "nsfem_advect": ([vp, C.c_int, vp, vp, flt, vp, vp], C.c_int),
    "nsfem_advect_f64": ([vp, C.c_int, vp, vp, dbl, vp, vp], C.c_int),
    "nsfem_locate": ([vp, C.c_int, vp, vp, vp, vp, vp], C.c_int),
    "nsfem_assemble": ([C.c_int, C.c_int, vp, C.c_int, vp, C.c_int, vp, vp, vp, C.c_int, vp, vp,
                        dbl, dbl, C.c_int, PP], C.c_int),
    "nsfem_assemble_device": ([vp, C.c_int, C.c_int, vp, C.c_int, vp, C.c_int, vp, vp, vp,
                               C.c_int, vp, vp, dbl, dbl, PP], C.c_int),
    "nsfem_assembly_destroy": ([vp], C.c_int),
    "nsfem_assembly_matrix_info": ([vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(i64)], C.c_int),
    "nsfem_assembly_matrix_copy": ([vp, C.c_int, vp, vp, vp], C.c_int),
    "nsfem_assembly_bc_copy": ([vp, vp, vp, vp], C.c_int),
    "nsfem_simulation_create": ([vp, C.c_int, C.c_int, vp, C.c_int, vp, C.c_int, vp, vp, vp,
                                 C.c_int, vp, vp, C.POINTER(SimOptions), PP], C.c_int),
    "nsfem_simulation_destroy": ([vp], C.c_int),
    "nsfem_simulation_step": ([vp, C.c_int, C.POINTER(StepStats)], C.c_int),
    "nsfem_simulation_get_fields": ([vp, vp, vp, vp], C.c_int),
    "nsfem_simulation_set_fields": ([vp, vp, vp, vp], C.c_int),
    "nsfem_simulation_set_options": ([vp, C.POINTER(SimOptions)], C.c_int),
    "nsfem_simulation_reset": ([vp], C.c_int),
    "nsfem_simulation_step_batch": ([vp, C.c_int, C.c_int, vp], C.c_int),
    "nsfem_simulation_autotune_batch": ([vp, C.c_int, C.POINTER(BatchTuning)], C.c_int),
    "nsfem_simulation_matrix_info": ([vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(i64)], C.c_int),
    "nsfem_simulation_sp_r_mult": ([vp, C.c_int, vp, vp, vp], C.c_int),
    "nsfem_simulation_set_warm_start": ([vp, C.c_int, vp], C.c_int),
    "nsfem_simulation_set_coarse_space": ([vp, C.c_int, C.c_int64, vp, vp, vp, C.c_double],
                                          C.c_int),
    "nsfem_simulation_add_coarse_level": ([vp, C.c_int, C.c_int64, vp, vp, vp], C.c_int),
    "nsfem_simulation_set_pressure_multigrid": ([vp, C.c_int], C.c_int),
    "nsfem_simulation_set_velocity_coarse_space": ([vp, C.c_int, C.c_int64, vp, vp, vp],
                                                   C.c_int),
    "nsfem_simulation_add_velocity_coarse_level": ([vp, C.c_int, C.c_int64, vp, vp, vp],
                                                   C.c_int),
    "nsfem_simulation_set_velocity_multigrid": ([vp, C.c_int], C.c_int),
    "nsfem_simulation_set_velocity_projection": ([vp, C.c_int], C.c_int),
    "nsfem_simulation_set_mass_projection": ([vp, C.c_int], C.c_int),
    "nsfem_simulation_copy_pressure_system": ([vp, vp, vp], C.c_int),
    "nsfem_gnn_create": ([vp, C.c_uint64, PP], C.c_int),
    "nsfem_gnn_destroy": ([vp], C.c_int),
    "nsfem_gnn_get_params": ([vp, vp], C.c_int),
    "nsfem_gnn_set_params": ([vp, vp], C.c_int),
    "nsfem_gnn_graph_create": ([vp, C.c_int, i64, vp, vp, vp, vp, vp, PP], C.c_int),
    "nsfem_gnn_graph_destroy": ([vp], C.c_int),
    "nsfem_gnn_predict": ([vp, vp, vp, vp], C.c_int),
    "nsfem_gnn_gradients": ([vp, vp, vp, vp, vp, C.POINTER(dbl)], C.c_int),
    "nsfem_gnn_train_step": ([vp, vp, vp, vp, flt, C.POINTER(dbl)], C.c_int),
}

EXPORTED = tuple(_SIGNATURES)


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP extension first "
                "(python -c 'import navier_stokes_fem_amd.build as b; b.build()')")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def last_error() -> str:
    return lib().nsfem_last_error().decode(errors="replace")


def check(status: int) -> None:
    if status != OK:
        raise NsfemError(status, last_error())


def ptr(a) -> int:
    """Address of a C-contiguous numpy array (None for None)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data
