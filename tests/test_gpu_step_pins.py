"""The device step's composition pinned by the same flows as the oracle's
(tests/step_cases.py): a uniform flow is a fixed point, a potential flow is
projected away with its pressure, a solid-body rotation stays put with the
centripetal pressure, the step is first order in dt, mirror symmetry holds
and the Dirichlet values are exact -- for the plain FP64 method, the FP32 +
FP64 refinement form, the chosen form (projection guess, coarse-grid
correction, velocity projection), the refinement form with
Jacobi-preconditioned velocity solves, the diffusion solved by V-cycle
preconditioned CG, and the whole chain of the 2048^2 headline (FP32
V-cycles for the diffusion and the pressure, Jacobi mass solves, projection
guesses of every solve) on the meshes built on a grid.  These do not go through the oracle: they
pin the device step (navier_stokes_fem_amd/csrc/simulation.cpp, after
include/assembly.h:1516-1554) directly.
"""
import numpy as np
import pytest

import step_cases as S

pytestmark = pytest.mark.gpu

FORMS = ("fp64", "fp32_refined", "chosen", "jacobi", "vcycle", "multigrid")


def gpu_stepper(dev, form, rtol=1e-10):
    from navier_stokes_fem_amd import mesh as M
    from navier_stokes_fem_amd.simulation import Simulation

    cache = {}

    def stepper(m, nu, dt, u, v, p, steps):
        key = (id(m), nu, dt)
        if key not in cache:
            for sim in cache.values():
                sim.close()
            cache.clear()
            sim = Simulation(m, nu, dt, device=dev, rtol=rtol,
                             precision="fp64" if form == "fp64" else "fp32_refined",
                             velocity_jacobi={"jacobi": True, "multigrid": "mass"}.get(form, False))
            if form in ("vcycle", "multigrid") and m.grid is not None:
                sim.set_velocity_multigrid(M.velocity_hierarchy(m, min_size=4), mode=2)
            if form == "multigrid" and m.grid is not None:
                H = M.grid_hierarchy(m, min_size=2)
                sim.set_coarse_space(H[0], rtol=1e-8)
                for level in H[1:]:
                    sim.add_coarse_level(level)
                sim.set_pressure_multigrid(True)
                sim.set_warm_start("projection")
                sim.set_velocity_projection(2)
                sim.set_mass_projection(1)
            if form == "chosen":
                sim.set_warm_start("projection")
                sim.set_velocity_projection(8)
                if m.grid is not None:
                    sim.set_coarse_space(M.grid_prolongation(m, 2), rtol=1e-8)
            cache[key] = sim
        sim = cache[key]
        sim.set_fields(u, v, p)
        st = sim.step(steps)
        assert max(st["max_relative_residual"]) <= rtol * (1 + 1e-6), st
        return sim.fields()

    stepper.close = lambda: [s.close() for s in cache.values()]
    return stepper


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("check", [S.check_uniform_flow, S.check_potential_flow,
                                   S.check_rotation, S.check_runge_order,
                                   S.check_mirror_symmetry, S.check_dirichlet_values],
                         ids=lambda f: f.__name__[6:])
def test_device_step_composition(gpu, form, check):
    print(f"device step pin {check.__name__} form {form}", flush=True)
    st = gpu_stepper(gpu, form)
    try:
        ok, detail = check(st)
    finally:
        st.close()
    assert ok, detail
