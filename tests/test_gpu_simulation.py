"""Whole time steps on the GPU against the oracle's steps (same inputs).

Bar: every CG solve reaches ||b - Ax|| <= 1e-8 ||b|| (true FP64 residual);
fields after N steps agree with the oracle's FP64 fields to 1e-6 relative
(the solves differ by at most cond * 1e-8 and advection is bit-exact given
equal inputs).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _oracle_run(oracle, m, nu, dt, steps, rtol):
    asm = oracle.Assembly(m, nu, dt)
    tree = oracle.kd_build(m.xy(), m.elems())
    u, v, p = oracle.initial_fields(m)
    its = []
    for _ in range(steps):
        u, v, p, st = oracle.step(asm, tree, m, dt, u, v, p, rtol=rtol)
        assert st["status"] == 0
        its.append(st["iterations"])
    return u, v, p, its


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("precision,pair", [("fp64", False), ("fp64", True),
                                            ("fp32_refined", True)])
def test_cavity32_steps(gpu, oracle, precision, pair):
    from navier_stokes_fem_amd import mesh
    from navier_stokes_fem_amd.simulation import Simulation

    m = mesh.cavity(32)
    steps = 5
    sim = Simulation(m, 0.01, 0.01, device=gpu, rtol=1e-8, precision=precision,
                     pair_velocity=pair)
    st = sim.step(steps)
    u, v, p = sim.fields()
    sim.close()
    ou, ov, op, its = _oracle_run(oracle, m, 0.01, 0.01, steps, 1e-8)
    assert max(st["max_relative_residual"]) <= 1e-8 * (1 + 1e-6)
    assert _rel(u, ou) < 1e-6 and _rel(v, ov) < 1e-6 and _rel(p, op) < 1e-6
    if precision == "fp64":
        tot = np.sum(its, axis=0)
        assert abs(st["cg_iterations"][2] - tot[2]) <= 2 * steps


def test_mesh_small_steps(gpu, oracle, mesh_small):
    from navier_stokes_fem_amd.simulation import Simulation

    m = mesh_small
    sim = Simulation(m, 0.001, 0.01, device=gpu, rtol=1e-8)
    st = sim.step(10)
    u, v, p = sim.fields()
    sim.close()
    ou, ov, op, _ = _oracle_run(oracle, m, 0.001, 0.01, 10, 1e-8)
    assert max(st["max_relative_residual"]) <= 1e-8 * (1 + 1e-6)
    assert _rel(u, ou) < 1e-6 and _rel(v, ov) < 1e-6 and _rel(p, op) < 1e-6


@pytest.mark.parametrize("precision", ["fp64", "fp32_refined"])
def test_projection_warm_start(gpu, precision):
    """The projection initial guess (A-norm projection onto an A-orthonormal
    basis of past pressure corrections) changes only where each pressure
    solve starts: the fields match the previous-pressure run, every solve
    meets the tolerance, and the pressure solves take fewer iterations.  40
    steps pass a basis restart (depth 16)."""
    from navier_stokes_fem_amd import mesh
    from navier_stokes_fem_amd.simulation import Simulation

    m = mesh.cavity(48)
    out = {}
    for mode in ("previous", "projection"):
        sim = Simulation(m, 1e-3, 1e-3, device=gpu, rtol=1e-8, precision=precision)
        sim.set_warm_start(mode)
        st = sim.step(40)
        out[mode] = (st, sim.fields())
        sim.close()
    (s0, f0), (s1, f1) = out["previous"], out["projection"]
    print(f"{precision}: pressure iterations previous {s0['cg_iterations'][2]}, "
          f"projection {s1['cg_iterations'][2]}")
    assert max(s1["max_relative_residual"]) <= 1e-8 * (1 + 1e-6)
    for a, b in zip(f1, f0):
        assert _rel(a, b) < 1e-6
    assert s1["cg_iterations"][2] < 0.9 * s0["cg_iterations"][2]


@pytest.mark.parametrize("precision,mode", [("fp64", "previous"), ("fp32_refined", "previous"),
                                            ("fp32_refined", "projection")])
def test_coarse_grid_correction(gpu, precision, mode):
    """The coarse-grid correction (P1 interpolation from the 64^2 cavity,
    A_c = P^T A P formed by the library) changes only where each pressure
    solve starts: the fields match the run without it, every solve meets the
    tolerance, the coarse solves ran, and the pressure solves take at least
    30% fewer iterations (about 50% on this cavity).  A wrong coarse operator
    (a dropped or transposed factor) loses that gain.  Removing the coarse
    space restores the plain guess."""
    from navier_stokes_fem_amd import mesh
    from navier_stokes_fem_amd.simulation import Simulation

    m = mesh.cavity(128)
    out = {}
    for coarse in (False, True):
        sim = Simulation(m, 1e-3, 1e-3, device=gpu, rtol=1e-8, precision=precision)
        sim.set_warm_start(mode)
        if coarse:
            sim.set_coarse_space(mesh.cavity_prolongation(128, 2), rtol=1e-6)
        sim.step(3)  # from rest: the first pressures are far from the later ones
        st = sim.step(12)
        out[coarse] = (st, sim.fields())
        if coarse:
            sim.set_coarse_space(None)
            assert sim.step(1)["coarse_iterations"] == 0
        sim.close()
    (s0, f0), (s1, f1) = out[False], out[True]
    print(f"{precision}/{mode}: pressure iterations {s0['cg_iterations'][2]} -> "
          f"{s1['cg_iterations'][2]}, coarse iterations {s1['coarse_iterations']}")
    assert max(s1["max_relative_residual"]) <= 1e-8 * (1 + 1e-6)
    for a, b in zip(f1, f0):
        assert _rel(a, b) < 1e-6
    assert s0["coarse_iterations"] == 0 and s1["coarse_iterations"] > 0
    assert s1["cg_iterations"][2] < 0.7 * s0["cg_iterations"][2]


@pytest.mark.parametrize("grids", [(64,), (64, 32, 16)])
def test_multigrid_coarse_levels(gpu, grids):
    """Multigrid levels below the coarse space (P1 interpolation 65^2 ->
    33^2, solved by CG on the deepest level, or 65^2 -> 33^2 -> 17^2 -> 9^2,
    the 9^2 level solved by its dense inverse, on the 128^2 cavity) change
    only how the coarse system is solved: V-cycles to the same coarse
    tolerance instead of CG.
    The fields match the run with the CG coarse solve, every pressure solve
    meets its tolerance, the pressure iteration counts stay within 10% of the
    CG coarse solve's, and the cycles number far fewer than the coarse CG's
    iterations.  A wrong level operator or transfer (a transposed P, a dropped
    smoothing step) stops the cycles from converging (no gain, or the
    iteration cap reached and the plain guess kept).  Setting the coarse
    space again removes the levels; a level before any coarse space, or one
    no smaller than the level above, is rejected."""
    from navier_stokes_fem_amd import mesh
    from navier_stokes_fem_amd.simulation import Simulation

    m = mesh.cavity(128)
    out = {}
    for levels in (False, True):
        sim = Simulation(m, 1e-3, 1e-3, device=gpu, rtol=1e-8, precision="fp32_refined")
        sim.set_warm_start("projection")
        if levels:
            with pytest.raises(ValueError):
                sim.add_coarse_level(mesh.cavity_prolongation(64, 2))
        sim.set_coarse_space(mesh.cavity_prolongation(128, 2), rtol=1e-6)
        if levels:
            for n in grids:
                sim.add_coarse_level(mesh.cavity_prolongation(n, 2))
        sim.step(3)
        st = sim.step(12)
        out[levels] = (st, sim.fields())
        if levels:
            k = (grids[-1] // 2 + 1) ** 2
            eye = (np.arange(k + 1), np.arange(k), np.ones(k), k)
            with pytest.raises(Exception):
                sim.add_coarse_level(eye)  # not smaller than the level above
            sim.set_coarse_space(mesh.cavity_prolongation(128, 2), rtol=1e-6)
            cg = sim.step(2)["coarse_iterations"]
            assert cg > 0
        sim.close()
    (s0, f0), (s1, f1) = out[False], out[True]
    print(f"pressure iterations {s0['cg_iterations'][2]} (CG coarse) -> "
          f"{s1['cg_iterations'][2]} (V-cycles); coarse {s0['coarse_iterations']} CG iterations, "
          f"{s1['coarse_iterations']} cycles")
    assert max(s1["max_relative_residual"]) <= 1e-8 * (1 + 1e-6)
    for a, b in zip(f1, f0):
        assert _rel(a, b) < 1e-6
    assert abs(s1["cg_iterations"][2] - s0["cg_iterations"][2]) <= 0.1 * s0["cg_iterations"][2]
    assert 0 < s1["coarse_iterations"] < 0.5 * s0["coarse_iterations"]


def test_vcycle_preconditioned_pressure_cg(gpu):
    """The pressure solve as CG preconditioned by one V-cycle (128^2 cavity,
    coarse space from 64^2, levels 32^2 and 16^2): every solve meets the
    tolerance, the fields match the plain CG run to 1e-6, the pressure
    solves take at most 15 iterations a step (the CPU premise: 7-9), no
    coarse correction runs, and a batch holding the simulation is rejected;
    switching it off restores the plain solve.  In the refined precision
    the cycles run in FP32 around the same FP64 CG."""
    from navier_stokes_fem_amd import mesh
    from navier_stokes_fem_amd.simulation import Simulation, step_batch

    m = mesh.cavity(128)
    out = {}
    for pmg in (False, True, "fp32"):
        sim = Simulation(m, 1e-3, 1e-3, device=gpu, rtol=1e-8,
                         precision="fp32_refined" if pmg == "fp32" else "fp64")
        if pmg:
            sim.set_coarse_space(mesh.cavity_prolongation(128, 2), rtol=1e-6)
            for n in (64, 32):
                sim.add_coarse_level(mesh.cavity_prolongation(n, 2))
            sim.set_pressure_multigrid(True)
        st = sim.step(8)
        out[pmg] = (st, sim.fields())
        if pmg is True:
            with pytest.raises(Exception):
                step_batch([sim], 1)
            sim.set_pressure_multigrid(False)
            assert sim.step(1)["cg_iterations"][2] > 15
        sim.close()
    (s0, f0), (s1, f1), (s2, f2) = out[False], out[True], out["fp32"]
    print(f"V-cycle preconditioned pressure CG: pressure iterations plain {s0['cg_iterations'][2]}, "
          f"V-cycle PCG {s1['cg_iterations'][2]}, FP32 cycles {s2['cg_iterations'][2]}", flush=True)
    for st in (s1, s2):
        assert max(st["max_relative_residual"]) <= 1e-8 * (1 + 1e-6)
        assert st["cg_iterations"][2] <= 15 * 8 and st["coarse_iterations"] == 0
        assert st["vcycle_solves"][2] == 8
    for a, b in zip(f1, f0):
        assert _rel(a, b) < 1e-6
    # the FP32 cycles (refined precision) precondition the same FP64 CG: the
    # pressure takes at most one more iteration a step and the fields match
    # as the mixed velocity solves allow
    assert s2["cg_iterations"][2] <= s1["cg_iterations"][2] + 8
    for a, b in zip(f2, f0):
        assert _rel(a, b) < 1e-5


def _errors_vs(fields, truth):
    return [float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(fields, truth)]


def _cavity_run(gpu, m, steps, precision, rtol, warm="previous", coarse=None, velocity=0,
                coarse_rtol=1e-6, levels=()):
    from navier_stokes_fem_amd.simulation import Simulation

    sim = Simulation(m, 1e-3, 1e-3, device=gpu, rtol=rtol, precision=precision)
    sim.set_warm_start(warm)
    if coarse is not None:
        sim.set_coarse_space(coarse, rtol=coarse_rtol)
        for lv in levels:
            sim.add_coarse_level(lv)
    sim.set_velocity_projection(velocity)
    st = sim.step(steps)
    f = sim.fields()
    sim.close()
    return f, st


def test_velocity_projection(gpu):
    """The diffusion guess u~ + projection of the correction: every solve
    meets the tolerance, the diffusion solves take far fewer iterations, and
    the fields are at least as close to a tightly solved run (FP64, 1e-13)
    as without it.  (Two guesses need not give the same fields to 1e-6:
    with the Dirichlet rows' O(1) values in ||b||, a 1e-8 relative residual
    fixes the interior velocity only to ~1e-4, the same for the reference's
    method.)  Depth 0 switches it off; a negative depth is rejected."""
    from navier_stokes_fem_amd import _native as N
    from navier_stokes_fem_amd import mesh
    from navier_stokes_fem_amd.simulation import Simulation

    m = mesh.cavity(64)
    truth, _ = _cavity_run(gpu, m, 10, "fp64", 1e-13)
    f0, s0 = _cavity_run(gpu, m, 10, "fp32_refined", 1e-8)
    f1, s1 = _cavity_run(gpu, m, 10, "fp32_refined", 1e-8, velocity=16)
    e0, e1 = _errors_vs(f0, truth), _errors_vs(f1, truth)
    d0 = s0["cg_iterations"][0] + s0["cg_iterations"][1]
    d1 = s1["cg_iterations"][0] + s1["cg_iterations"][1]
    print(f"diffusion iterations {d0} -> {d1}; errors vs 1e-13 run {e0} -> {e1}")
    assert max(s1["max_relative_residual"]) <= 1e-8 * (1 + 1e-6)
    assert d1 < 0.7 * d0
    for a, b in zip(e1, e0):
        assert a <= 1.2 * b
    sim = Simulation(m, 1e-3, 1e-3, device=gpu)
    sim.set_velocity_projection(16)
    sim.step(2)
    sim.set_velocity_projection(0)
    sim.step(1)
    with pytest.raises(N.NsfemError):
        sim.set_velocity_projection(-1)
    sim.close()


@pytest.mark.parametrize("chain", ["coarse_cg", "coarse_cg_velocity16", "bench_chain"])
def test_chosen_form_is_as_accurate_as_the_plain_method(gpu, chain):
    """FP32 + FP64 refinement, projection guess and coarse correction (with
    and without the velocity projection) against the plain FP64 method with
    the previous guesses, both at 1e-8: per field, the chosen form's distance
    to a tightly solved run (FP64, 1e-13) is at most 1.2x the plain
    method's.  ``bench_chain`` is the headline's configuration on the 128^2
    cavity: coarse space from the n/2 cavity, its system by V-cycles over
    levels down to 16^2 to a coarse tolerance of 1e-5, velocity projection
    depth 8."""
    from navier_stokes_fem_amd import mesh

    m = mesh.cavity(128)
    truth, _ = _cavity_run(gpu, m, 10, "fp64", 1e-13)
    fp, _ = _cavity_run(gpu, m, 10, "fp64", 1e-8)
    if chain == "bench_chain":
        fc, sc = _cavity_run(gpu, m, 10, "fp32_refined", 1e-8, warm="projection",
                             coarse=mesh.cavity_prolongation(128, 2), velocity=8,
                             coarse_rtol=1e-5,
                             levels=[mesh.cavity_prolongation(n, 2) for n in (64, 32)])
        assert sc["coarse_iterations"] > 0
    else:
        fc, sc = _cavity_run(gpu, m, 10, "fp32_refined", 1e-8, warm="projection",
                             coarse=mesh.cavity_prolongation(128, 4),
                             velocity=16 if chain.endswith("16") else 0)
    ep, ec = _errors_vs(fp, truth), _errors_vs(fc, truth)
    print(f"rel L2 error vs 1e-13 run (u, v, p): plain {ep}, chosen {ec}")
    assert max(sc["max_relative_residual"]) <= 1e-8 * (1 + 1e-6)
    for a, b in zip(ec, ep):
        assert a <= 1.2 * b


def test_coarse_space_rejects_bad_input(gpu):
    from navier_stokes_fem_amd import _native as N
    from navier_stokes_fem_amd import mesh
    from navier_stokes_fem_amd.simulation import Simulation

    m = mesh.cavity(8)
    sim = Simulation(m, 1e-3, 1e-3, device=gpu)
    rs, col, val, nc = mesh.cavity_prolongation(8, 2)
    with pytest.raises(N.NsfemError):
        sim.set_coarse_space((rs, col, val, nc - 1))  # a column out of range
    with pytest.raises(N.NsfemError):
        sim.set_coarse_space((rs, col, val, nc), rtol=0.0)
    bad = col.copy()
    bad[[1, 2]] = bad[[2, 1]]  # columns out of order within a row
    with pytest.raises(N.NsfemError):
        sim.set_coarse_space((rs, bad, val, nc))
    assert sim.step(1)["coarse_iterations"] == 0  # nothing was set
    sim.close()


def test_failed_coarse_space_leaves_the_plain_solve(gpu):
    """A coarse space whose set-up fails part way (here a P column with no
    entries: a zero coarse diagonal) leaves no coarse space and no V-cycle
    preconditioner behind: the next steps are the plain solve's, bit for
    bit."""
    from navier_stokes_fem_amd import _native as N
    from navier_stokes_fem_amd import mesh
    from navier_stokes_fem_amd.simulation import Simulation

    m = mesh.cavity(16)
    sim = Simulation(m, 1e-3, 1e-3, device=gpu, rtol=1e-8)
    rs, col, val, nc = mesh.cavity_prolongation(16, 2)
    sim.set_coarse_space((rs, col, val, nc))
    sim.set_pressure_multigrid(True)
    with pytest.raises(N.NsfemError):
        sim.set_coarse_space((rs, col, val, nc + 1))  # the last column is empty
    st = sim.step(2)
    f = sim.fields()
    sim.close()
    ref = Simulation(m, 1e-3, 1e-3, device=gpu, rtol=1e-8)
    st0 = ref.step(2)
    f0 = ref.fields()
    ref.close()
    assert st["coarse_iterations"] == 0
    assert st["cg_iterations"] == st0["cg_iterations"]
    for a, b in zip(f, f0):
        np.testing.assert_array_equal(a, b)


def test_device_closed_under_a_simulation():
    """Closing the device releases a simulation's memory, its coarse space
    and projection bases included; every later call on the simulation
    raises instead of touching freed memory, and closing it still works."""
    from navier_stokes_fem_amd import _native as N
    from navier_stokes_fem_amd import mesh
    from navier_stokes_fem_amd.device import GPUSimulationDevice
    from navier_stokes_fem_amd.simulation import Simulation

    dev = GPUSimulationDevice(0)
    sim = Simulation(mesh.cavity(16), 1e-3, 1e-3, device=dev, precision="fp32_refined")
    sim.set_warm_start("projection")
    sim.set_coarse_space(mesh.cavity_prolongation(16, 2))
    sim.set_velocity_projection(4)
    assert sim.step(2)["coarse_iterations"] > 0
    dev.close()
    for call in (lambda: sim.step(1), lambda: sim.set_coarse_space(None),
                 lambda: sim.set_velocity_projection(0), lambda: sim.fields()):
        with pytest.raises(N.NsfemError):
            call()
    sim.close()


def test_export_fields(gpu, tmp_path):
    """The exported result file holds exactly the simulation's fields."""
    from navier_stokes_fem_amd import mesh, results
    from navier_stokes_fem_amd.simulation import Simulation

    m = mesh.cavity(16)
    sim = Simulation(m, 0.01, 0.01, device=gpu, rtol=1e-8)
    sim.step(3)
    path = sim.export(str(tmp_path), 3)
    u, v, p = sim.fields()
    sim.close()
    sol = results.read_solution(path)
    assert sol["timeStep"] == 3
    np.testing.assert_array_equal(sol["u"], u)
    np.testing.assert_array_equal(sol["v"], v)
    np.testing.assert_array_equal(sol["p"], p)

def test_velocity_jacobi(gpu):
    """Jacobi-preconditioned velocity solves (CG on D^-1/2 A D^-1/2 in FP32,
    x, b and the FP64 residual passes unscaled): every solve meets the same
    unscaled 1e-8 relative residual, the diffusion and mass solves take
    fewer iterations than unscaled, the pressure solve is untouched, and the
    fields are as close to a tightly solved FP64 run as the unscaled form's."""
    from navier_stokes_fem_amd import mesh
    from navier_stokes_fem_amd.simulation import Simulation

    m = mesh.cavity(96)
    out = {}
    for form, prec, jac, rtol in (("tight", "fp64", False, 1e-12), ("plain", "fp32_refined", False, 1e-8),
                                  ("jacobi", "fp32_refined", True, 1e-8)):
        sim = Simulation(m, 1e-3, 1e-3, device=gpu, rtol=rtol, precision=prec, velocity_jacobi=jac,
                         pair_velocity=False)
        st = sim.step(10)
        out[form] = (st, sim.fields())
        sim.close()
    (sp_, fp), (sj, fj), (_, ft) = out["plain"], out["jacobi"], out["tight"]
    print("velocity iterations plain", sp_["cg_iterations"], "jacobi", sj["cg_iterations"])
    assert max(sj["max_relative_residual"]) <= 1e-8 * (1 + 1e-6)
    for k in (0, 1, 3, 4):
        assert sj["cg_iterations"][k] <= sp_["cg_iterations"][k], k
    assert sum(sj["cg_iterations"][k] for k in (1, 3, 4)) < 0.95 * sum(
        sp_["cg_iterations"][k] for k in (1, 3, 4))
    for a, b, t in zip(fj, fp, ft):
        ej, ep = _rel(a, t), _rel(b, t)
        print(f"distance to the 1e-12 run: jacobi {ej:.3e}, plain {ep:.3e}")
        assert ej <= 2.0 * ep + 1e-7


def test_vcycle_preconditioned_diffusion_cg(gpu):
    """The diffusion columns as CG preconditioned by one V-cycle over the
    velocity nodes, the P1 vertex space (Dirichlet vertices left out) and
    one halving, solved by its dense inverse (64^2 cavity, nu dt / h^2 as
    on the 2048^2 cavity at
    nu = dt = 1e-3): every solve meets the tolerance, the fields match the
    plain CG's to 1e-7 (both at 1e-11), the diffusion columns take at most
    15 iterations a step against more than 3x that plainly, the pressure and
    mass solves are untouched, and the stats count the V-cycle solves; the
    mixed precision's diffusion solves run the same FP64 preconditioned CG."""
    from navier_stokes_fem_amd import mesh
    from navier_stokes_fem_amd.simulation import Simulation

    m = mesh.cavity(64)
    nu = 1e-3 * (2048 / 64) ** 2
    H = mesh.velocity_hierarchy(m)
    assert [h[3] for h in H] == [63 * 63, 31 * 31]
    out = {}
    for form, prec, mode in (("plain", "fp64", None), ("vcycle", "fp64", 2),
                             ("mixed", "fp32_refined", 2)):
        sim = Simulation(m, nu, 1e-3, device=gpu, rtol=1e-11, precision=prec)
        if mode is not None:
            sim.set_velocity_multigrid(H, mode=mode)
        st = sim.step(4)
        out[form] = (st, sim.fields())
        sim.close()
    (s0, f0), (s1, f1), (s2, f2) = out["plain"], out["vcycle"], out["mixed"]
    print("diffusion iterations plain", s0["cg_iterations"][:2], "V-cycle", s1["cg_iterations"][:2],
          "mixed", s2["cg_iterations"][:2], flush=True)
    for st in (s1, s2):
        assert max(st["max_relative_residual"]) <= 1e-11 * (1 + 1e-6)
        assert st["vcycle_solves"][:2] == [4, 4] and st["vcycle_solves"][2:] == [0, 0, 0]
        assert max(st["cg_iterations"][:2]) <= 15 * 4
    assert s0["vcycle_solves"] == [0] * 5
    assert s0["cg_iterations"][1] > 3 * s1["cg_iterations"][1]
    for a, b in zip(f1, f0):
        assert _rel(a, b) < 1e-7
    for a, b in zip(f2, f0):
        assert _rel(a, b) < 1e-7


def test_velocity_multigrid_per_column_choice_and_errors(gpu):
    """Mode 1 times both ways per column (the V-cycle first, the plain CG
    next) and then takes the faster; a hierarchy whose first prolongation
    has a row on a Dirichlet node, or a mode without a hierarchy, is
    rejected and leaves the plain solve; a batch holding the simulation is
    rejected; removing the hierarchy restores the plain solves."""
    from navier_stokes_fem_amd import _native as N
    from navier_stokes_fem_amd import mesh
    from navier_stokes_fem_amd.simulation import Simulation, step_batch

    m = mesh.cavity(48)
    nu = 1e-3 * (2048 / 48) ** 2
    H = mesh.velocity_hierarchy(m)
    sim = Simulation(m, nu, 1e-3, device=gpu, rtol=1e-8)
    with pytest.raises(N.NsfemError):
        sim.set_velocity_multigrid(None, mode=1)  # no hierarchy yet
    rs, col, val, nc = H[0]
    d = int(m.vel_bc_nodes[0])
    bad_rs = rs.copy()
    bad_rs[d + 1:] += 1
    bad_col = np.insert(col, rs[d], 0)
    bad_val = np.insert(val, rs[d], 1.0)
    with pytest.raises(N.NsfemError, match="Dirichlet"):
        sim.set_velocity_multigrid([(bad_rs, bad_col, bad_val, nc)], mode=1)
    with pytest.raises(N.NsfemError):
        sim.set_velocity_multigrid(None, mode=2)  # the failed call left none
    sim.set_velocity_multigrid(H, mode=1)
    with pytest.raises(N.NsfemError):
        step_batch([sim], 1)
    first = [sim.step(1)["vcycle_solves"][:2] for _ in range(4)]
    print("V-cycle solves per step, mode 1:", first, flush=True)
    assert first[0] == [1, 1] and first[1] == [0, 0]
    st = sim.step(2)
    assert max(st["max_relative_residual"]) <= 1e-8 * (1 + 1e-6)
    sim.set_velocity_multigrid(None, mode=0)  # hierarchy kept, plain solves
    assert sim.step(1)["vcycle_solves"] == [0] * 5
    sim.set_velocity_multigrid(None, mode=2)
    assert sim.step(1)["vcycle_solves"][:2] == [1, 1]
    sim.remove_velocity_multigrid()
    st = sim.step(1)
    assert st["vcycle_solves"] == [0] * 5
    assert max(st["max_relative_residual"]) <= 1e-8 * (1 + 1e-6)
    with pytest.raises(N.NsfemError):
        sim.set_velocity_multigrid(None, mode=1)
    sim.close()


def test_mass_projection(gpu):
    """The mass solves started from the A-norm projection of the velocity
    correction onto its past solutions: every solve meets the tolerance,
    the mass solves take fewer iterations than from the reference's guess
    (the velocity), the other solves are untouched in kind, and the fields
    are as close to a tightly solved run; depth 0 restores the reference's
    guess, a negative depth is rejected and a batch holding the simulation
    is rejected."""
    from navier_stokes_fem_amd import _native as N
    from navier_stokes_fem_amd import mesh
    from navier_stokes_fem_amd.simulation import Simulation, step_batch

    m = mesh.cavity(64)
    out = {}
    for form, depth, rtol, prec in (("tight", 0, 1e-12, "fp64"), ("plain", 0, 1e-8, "fp32_refined"),
                                    ("projection", 4, 1e-8, "fp32_refined")):
        sim = Simulation(m, 1e-3, 1e-3, device=gpu, rtol=rtol, precision=prec)
        sim.set_mass_projection(depth)
        st = sim.step(10)
        out[form] = (st, sim.fields())
        if depth:
            with pytest.raises(N.NsfemError):
                step_batch([sim], 1)
            with pytest.raises(N.NsfemError):
                sim.set_mass_projection(-1)
            sim.set_mass_projection(0)
            assert max(sim.step(1)["max_relative_residual"]) <= 1e-8 * (1 + 1e-6)
        sim.close()
    (s0, f0), (s1, f1), (_, ft) = out["plain"], out["projection"], out["tight"]
    print("mass iterations reference guess", s0["cg_iterations"][3:], "projection",
          s1["cg_iterations"][3:], flush=True)
    assert max(s1["max_relative_residual"]) <= 1e-8 * (1 + 1e-6)
    assert sum(s1["cg_iterations"][3:]) < 0.85 * sum(s0["cg_iterations"][3:])
    for a, b, t in zip(f1, f0, ft):
        assert _rel(a, t) <= 2.0 * _rel(b, t) + 1e-7


def test_captured_vcycles_replay_like_eager_launches(gpu, monkeypatch):
    """With NSFEM_VCYCLE_GRAPH=1 the FP32 V-cycles of the pressure and
    diffusion solves run from a HIP graph captured on their first use and
    replayed on every later one, with the CG's own launches, dots and host
    synchronisations between the replays: 6 steps (dozens of replays of
    each graph) give bit-identical fields and iteration counts to the same
    steps with every cycle launched directly (the default), twice over."""
    from navier_stokes_fem_amd import mesh
    from navier_stokes_fem_amd.simulation import Simulation

    m = mesh.cavity(64)
    nu = 1e-3 * (2048 / 64) ** 2
    out = {}
    for graph in ("1", "0", "1"):
        monkeypatch.setenv("NSFEM_VCYCLE_GRAPH", graph)
        # pairing off: its choice follows measured times, which differ run to run
        sim = Simulation(m, nu, 1e-3, device=gpu, rtol=1e-8, precision="fp32_refined",
                         velocity_jacobi="mass", pair_velocity=False)
        H = mesh.grid_hierarchy(m, min_size=8)
        sim.set_coarse_space(H[0], rtol=1e-6)
        for level in H[1:]:
            sim.add_coarse_level(level)
        sim.set_pressure_multigrid(True)
        sim.set_velocity_multigrid(mesh.velocity_hierarchy(m, min_size=8), mode=2)
        st = sim.step(6)
        out.setdefault(graph, []).append((st["cg_iterations"], sim.fields()))
        assert max(st["max_relative_residual"]) <= 1e-8 * (1 + 1e-6)
        assert st["vcycle_solves"][:3] == [6, 6, 6]
        sim.close()
    (it_g, f_g), (it_g2, f_g2) = out["1"]
    ((it_e, f_e),) = out["0"]
    print("iterations with graphs", it_g, "eager", it_e, flush=True)
    assert it_g == it_e == it_g2
    for a, b, c in zip(f_g, f_e, f_g2):
        np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(a, c)
