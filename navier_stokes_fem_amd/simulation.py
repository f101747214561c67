"""Device-resident semi-Lagrangian Navier-Stokes stepping
(NavierStokesAssembly::semiLagrangianSolve, include/assembly.h:1477-1557).

``Simulation(mesh, viscosity, dt)`` assembles the matrices on the host,
builds the KD tree, uploads everything and keeps ``u, v, p`` in HBM; each
:meth:`Simulation.step` runs advect + 5 CG solves per step on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .device import GPUSimulationDevice

__all__ = ["Simulation", "step_batch", "autotune_batch"]


class Simulation:
    def __init__(self, mesh, viscosity: float, dt: float, device: GPUSimulationDevice = None,
                 rtol: float = 1e-8, precision: str = "fp64", pair_velocity: bool = True,
                 max_iterations: int = -1, threads: int = 0, assembly: str = "device",
                 velocity_jacobi: bool = False):
        self.mesh = mesh
        self.device = device if device is not None else GPUSimulationDevice(0)
        self.options = N.SimOptions(viscosity=float(viscosity), dt=float(dt), rtol=float(rtol),
                                    max_iterations=int(max_iterations),
                                    precision={"fp64": N.PRECISION_FP64,
                                               "fp32_refined": N.PRECISION_FP32_REFINED}[precision],
                                    pair_velocity=int(bool(pair_velocity)),
                                    threads=int(threads),
                                    assembly={"device": N.ASSEMBLY_DEVICE,
                                              "host": N.ASSEMBLY_HOST}[assembly],
                                    velocity_jacobi=self._jacobi_bits(velocity_jacobi))
        xy, el = mesh.xy(), mesh.elems()
        vb = np.ascontiguousarray(mesh.vel_bc_nodes, np.int32)
        gu = np.ascontiguousarray(mesh.vel_bc_u, np.float64)
        gv = np.ascontiguousarray(mesh.vel_bc_v, np.float64)
        pb = np.ascontiguousarray(mesh.p_bc_nodes, np.int32)
        gp = np.ascontiguousarray(mesh.p_bc_values, np.float64)
        h = C.c_void_p()
        N.check(N.lib().nsfem_simulation_create(
            self.device.handle, mesh.node_count, mesh.pressure_count, N.ptr(xy),
            mesh.element_count, N.ptr(el), len(vb), N.ptr(vb) if len(vb) else None,
            N.ptr(gu) if len(vb) else None, N.ptr(gv) if len(vb) else None, len(pb),
            N.ptr(pb) if len(pb) else None, N.ptr(gp) if len(pb) else None,
            C.byref(self.options), C.byref(h)))
        self.handle = h

    def step(self, steps: int = 1) -> dict:
        st = N.StepStats()
        N.check(N.lib().nsfem_simulation_step(self.handle, int(steps), C.byref(st)))
        return st.as_dict()

    def fields(self):
        V, P = self.mesh.node_count, self.mesh.pressure_count
        u, v, p = np.empty(V), np.empty(V), np.empty(P)
        N.check(N.lib().nsfem_simulation_get_fields(self.handle, N.ptr(u), N.ptr(v), N.ptr(p)))
        return u, v, p

    def set_fields(self, u=None, v=None, p=None):
        a = [None if x is None else np.ascontiguousarray(x, np.float64) for x in (u, v, p)]
        N.check(N.lib().nsfem_simulation_set_fields(self.handle, *[N.ptr(x) for x in a]))

    @staticmethod
    def _jacobi_bits(v) -> int:
        """``velocity_jacobi``: False / "none", "mass", "diffusion", True /
        "both" -> the option's bits (1 = mass, 2 = diffusion)."""
        if isinstance(v, str):
            return {"none": 0, "mass": 1, "diffusion": 2, "both": 3}[v]
        return 3 if v else 0

    def configure(self, rtol=None, precision=None, pair_velocity=None, max_iterations=None,
                  velocity_jacobi=None):
        """Change solver options (viscosity and dt are fixed at creation).
        ``velocity_jacobi``: in the refined precision, the mass and/or
        diffusion solves as Jacobi-preconditioned CG (same stopping rule):
        False / "none", "mass", "diffusion", True / "both"."""
        o = self.options
        if velocity_jacobi is not None:
            o.velocity_jacobi = self._jacobi_bits(velocity_jacobi)
        if rtol is not None:
            o.rtol = float(rtol)
        if precision is not None:
            o.precision = {"fp64": N.PRECISION_FP64, "fp32_refined": N.PRECISION_FP32_REFINED}[precision]
        if pair_velocity is not None:
            o.pair_velocity = int(bool(pair_velocity))
        if max_iterations is not None:
            o.max_iterations = int(max_iterations)
        N.check(N.lib().nsfem_simulation_set_options(self.handle, C.byref(o)))

    def reset(self):
        """Zero fields with the velocity Dirichlet values imposed."""
        N.check(N.lib().nsfem_simulation_reset(self.handle))

    def matrix_info(self, matrix: int):
        """(rows, cols, nnz) of one of the simulation's matrices (ids as the
        device slots: 0 pressure stiffness, 1 velocity mass, 2 diffusion,
        3 velocity divergence, 4 pressure divergence)."""
        r, c, z = C.c_int(0), C.c_int(0), C.c_int64(0)
        N.check(N.lib().nsfem_simulation_matrix_info(self.handle, int(matrix), C.byref(r),
                                                       C.byref(c), C.byref(z)))
        return r.value, c.value, z.value

    def sp_r_mult(self, matrix: int, x, y, lhs=None):
        """y = A x, or y = lhs - A x, with one of the simulation's matrices."""
        N.check(N.lib().nsfem_simulation_sp_r_mult(self.handle, int(matrix),
                                                   lhs.handle if lhs is not None else None,
                                                   x.handle, y.handle))

    WARM_MODES = {"previous": N.WARM_PREVIOUS, "zero": N.WARM_ZERO, "learned": N.WARM_LEARNED,
                  "learned_or_previous": N.WARM_LEARNED_OR_PREVIOUS,
                  "learned_correction": N.WARM_LEARNED_CORRECTION,
                  "extrapolate": N.WARM_EXTRAPOLATE, "projection": N.WARM_PROJECTION,
                  "learned_search": N.WARM_LEARNED_SEARCH,
                  "learned_projection": N.WARM_LEARNED_PROJECTION}

    def set_warm_start(self, mode: str = "previous", net=None):
        """Initial guess of the pressure solve: ``previous`` (the reference's
        choice), ``zero``, ``learned`` (``net``'s prediction from the
        right-hand side), ``learned_or_previous`` (the smaller residual) or
        ``learned_correction`` (previous pressure + the network's correction
        from the previous pressure's residual, if that lowers the residual),
        ``extrapolate`` (2 p_n - p_(n-1), if that lowers the residual),
        ``projection`` (the A-norm projection of the solution onto the span of
        past pressures, Fischer 1998), ``learned_search`` (previous pressure
        plus the network's correction taken with the A-norm-optimal step:
        never worse than the previous pressure in the A-norm) or
        ``learned_projection`` (the same step from the projection guess)."""
        N.check(N.lib().nsfem_simulation_set_warm_start(
            self.handle, self.WARM_MODES[mode], net.handle if net is not None else None))

    def export(self, folder: str, time_step: int):
        """Write the current fields as ``<folder>/out_<time_step>.json``, the
        reference's result file (include/assembly.h:953-979)."""
        from .results import export_simulation

        return export_simulation(self, folder, time_step)

    def set_velocity_projection(self, depth: int = 16):
        """Start each diffusion solve from the advected velocity plus the
        A-norm projection of its correction onto the last ``depth``
        corrections (0 = the advected velocity alone, the reference's)."""
        N.check(N.lib().nsfem_simulation_set_velocity_projection(self.handle, int(depth)))

    def set_mass_projection(self, depth: int = 2):
        """Start each mass solve (the velocity correction) from the A-norm
        projection onto its last ``depth`` solutions instead of the
        reference's guess, the velocity itself (0 = the reference's)."""
        N.check(N.lib().nsfem_simulation_set_mass_projection(self.handle, int(depth)))

    def set_coarse_space(self, prolongation=None, rtol: float = 1e-6):
        """Coarse-grid correction of every pressure solve's initial guess:
        ``prolongation`` = (row_start, col, val, coarse_count), one row per
        pressure node (e.g. ``mesh.cavity_prolongation``); the library forms
        P^T A P and solves the restricted residual equation to ``rtol`` before
        each pressure CG.  None removes it."""
        self._coarse_sizes = []
        if prolongation is None:
            N.check(N.lib().nsfem_simulation_set_coarse_space(self.handle, 0, 0, None, None, None,
                                                              float(rtol)))
            return
        rs, col, val, nc = prolongation
        rs = np.ascontiguousarray(rs, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        val = np.ascontiguousarray(val, np.float64)
        if len(rs) != self.mesh.pressure_count + 1:
            raise ValueError("the prolongation needs one row per pressure node")
        N.check(N.lib().nsfem_simulation_set_coarse_space(self.handle, int(nc), len(col),
                                                          N.ptr(rs), N.ptr(col), N.ptr(val),
                                                          float(rtol)))
        self._coarse_sizes = [int(nc)]

    def add_coarse_level(self, prolongation):
        """Add a multigrid level below the deepest coarse level:
        ``prolongation`` = (row_start, col, val, count) with one row per node
        of the level above (e.g. ``mesh.cavity_prolongation(n_above, 2)``).
        With levels the coarse system is solved by V-cycles instead of CG."""
        rs, col, val, nc = prolongation
        rs = np.ascontiguousarray(rs, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        val = np.ascontiguousarray(val, np.float64)
        sizes = getattr(self, "_coarse_sizes", [])
        if not sizes:
            raise ValueError("set a coarse space before adding levels")
        if len(rs) != sizes[-1] + 1:
            raise ValueError("the level's prolongation needs one row per node of the level above")
        N.check(N.lib().nsfem_simulation_add_coarse_level(self.handle, int(nc), len(col),
                                                          N.ptr(rs), N.ptr(col), N.ptr(val)))
        sizes.append(int(nc))

    def set_pressure_multigrid(self, enable: bool = True):
        """Solve the pressure by CG preconditioned with one V-cycle over the
        fine grid, the coarse space and its levels (FP64; needs a coarse
        space; single-simulation stepping only)."""
        N.check(N.lib().nsfem_simulation_set_pressure_multigrid(self.handle, int(bool(enable))))

    def set_velocity_multigrid(self, hierarchy=None, mode: int = 1):
        """Solve the diffusion columns by CG preconditioned with one V-cycle
        over the velocity nodes and ``hierarchy`` (a list of prolongations
        (row_start, col, val, count): the first with one row per velocity
        node, zero rows on the velocity Dirichlet nodes, each next one with a
        row per node of the level above, e.g. ``mesh.velocity_hierarchy``);
        the library forms the Galerkin matrices.  ``mode``: 1 = each column
        the way (plain CG or V-cycle) measured faster, 2 = both columns by
        V-cycle, 0 = keep the hierarchy but solve plainly.  ``hierarchy=None``
        keeps the hierarchy set before and changes the mode only."""
        lib = N.lib()
        if hierarchy is None:
            N.check(lib.nsfem_simulation_set_velocity_multigrid(self.handle, int(mode)))
            return
        above = self.mesh.node_count
        self._velocity_sizes = []
        for k, (rs, col, val, nc) in enumerate(hierarchy):
            rs = np.ascontiguousarray(rs, np.int32)
            col = np.ascontiguousarray(col, np.int32)
            val = np.ascontiguousarray(val, np.float64)
            if len(rs) != above + 1:
                raise ValueError(f"prolongation {k} needs one row per node of the level above "
                                 f"({above})")
            fn = (lib.nsfem_simulation_set_velocity_coarse_space if k == 0 else
                  lib.nsfem_simulation_add_velocity_coarse_level)
            N.check(fn(self.handle, int(nc), len(col), N.ptr(rs), N.ptr(col), N.ptr(val)))
            self._velocity_sizes.append(int(nc))
            above = int(nc)
        N.check(lib.nsfem_simulation_set_velocity_multigrid(self.handle, int(mode)))

    def remove_velocity_multigrid(self):
        """Drop the velocity hierarchy: the diffusion solves are plain again."""
        N.check(N.lib().nsfem_simulation_set_velocity_coarse_space(self.handle, 0, 0, None, None,
                                                                   None))
        self._velocity_sizes = []

    def copy_pressure_system(self, rhs=None, p=None):
        """Copy the last step's pressure right-hand side and solution into
        device buffers (``DeviceBuffer`` of pressure-count doubles)."""
        N.check(N.lib().nsfem_simulation_copy_pressure_system(
            self.handle, rhs.handle if rhs is not None else None,
            p.handle if p is not None else None))

    def close(self):
        if getattr(self, "handle", None):
            N.lib().nsfem_simulation_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _handles(sims):
    arr = (C.c_void_p * len(sims))(*[s.handle.value if hasattr(s.handle, "value") else s.handle
                                     for s in sims])
    return arr


def step_batch(sims, steps: int = 1) -> list:
    """Advance several simulations (one device, one precision) together:
    each solve kind runs for all of them in one persistent launch.  Returns
    one stats dict per simulation (seconds are the whole batch's)."""
    stats = (N.StepStats * len(sims))()
    N.check(N.lib().nsfem_simulation_step_batch(_handles(sims), len(sims), int(steps),
                                                C.cast(stats, C.c_void_p)))
    return [s.as_dict() for s in stats]


def autotune_batch(sims) -> dict:
    """Pick the wave-tile shape and blocks per instance of the batched solves
    by timing the candidates; applied to every matrix of the batch."""
    t = N.BatchTuning()
    N.check(N.lib().nsfem_simulation_autotune_batch(_handles(sims), len(sims), C.byref(t)))
    return t.as_dict()
