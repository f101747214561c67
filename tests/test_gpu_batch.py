"""Batched stepping (several simulations, each solve kind in one persistent
launch for all of them) against stepping the same simulations one by one,
and the autotuner of the batched solves."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(gpu, specs, precision, rtol):
    from navier_stokes_fem_amd import mesh
    from navier_stokes_fem_amd.simulation import Simulation

    sims = []
    for n, lid, nu in specs:
        sims.append(Simulation(mesh.cavity(n, lid_speed=lid), nu, 0.005, device=gpu, rtol=rtol,
                               precision=precision))
    return sims


SPECS = [(8, 1.0, 0.01), (20, 0.7, 0.002), (12, 1.9, 0.005), (16, 1.2, 0.001), (10, 0.5, 0.008)]


@pytest.mark.parametrize("precision", ["fp64", "fp32_refined"])
def test_batch_matches_individual_steps(gpu, precision):
    from navier_stokes_fem_amd.simulation import step_batch

    # the FP32 form is run at 1e-9: on these tiny systems a tighter
    # tolerance exhausts the default iteration limit (the row count)
    rtol = 1e-11 if precision == "fp64" else 1e-9
    tol = 1e-8 if precision == "fp64" else 1e-6
    one = _batch(gpu, SPECS, precision, rtol)
    many = _batch(gpu, SPECS, precision, rtol)
    for s in one:
        s.step(4)
    stats = step_batch(many, 4)
    assert len(stats) == len(SPECS)
    for a, b, st in zip(one, many, stats):
        assert st["steps"] == 4
        assert max(st["max_relative_residual"]) <= rtol * 1.01
        for fa, fb in zip(a.fields(), b.fields()):
            np.testing.assert_allclose(fb, fa, rtol=0, atol=tol * max(1.0, np.abs(fa).max()))
    for s in one + many:
        s.close()


def test_batch_autotune_keeps_results(gpu):
    from navier_stokes_fem_amd.simulation import autotune_batch, step_batch

    ref = _batch(gpu, SPECS, "fp64", 1e-11)
    tuned = _batch(gpu, SPECS, "fp64", 1e-11)
    t = autotune_batch(tuned)
    print("autotune:", t)
    assert t["rows_per_wave"] in (64, 32, 16) and t["block_scale"] in (1, 2, 4)
    assert t["velocity_block_scale"] in (1, 2, 4)
    assert t["seconds_per_iteration"] > 0
    step_batch(ref, 3)
    step_batch(tuned, 3)
    for a, b in zip(ref, tuned):
        for fa, fb in zip(a.fields(), b.fields()):
            np.testing.assert_allclose(fb, fa, rtol=0, atol=1e-8 * max(1.0, np.abs(fa).max()))
    for s in ref + tuned:
        s.close()


def test_batch_of_many_small_instances(gpu):
    """More instances than a block each would need: 40 cavities of 4..12."""
    from navier_stokes_fem_amd.simulation import step_batch

    rng = np.random.default_rng(5)
    specs = [(int(rng.integers(4, 13)), float(rng.uniform(0.5, 2)), 0.01) for _ in range(40)]
    sims = _batch(gpu, specs, "fp64", 1e-10)
    stats = step_batch(sims, 2)
    for st in stats:
        assert max(st["max_relative_residual"]) <= 1e-10 * 1.01
    # one of them stepped alone gives the same fields
    from navier_stokes_fem_amd import mesh
    from navier_stokes_fem_amd.simulation import Simulation

    n, lid, nu = specs[7]
    alone = Simulation(mesh.cavity(n, lid_speed=lid), nu, 0.005, device=gpu, rtol=1e-10)
    alone.step(2)
    for fa, fb in zip(alone.fields(), sims[7].fields()):
        np.testing.assert_allclose(fb, fa, rtol=0, atol=1e-7 * max(1.0, np.abs(fa).max()))
    alone.close()
    for s in sims:
        s.close()


@pytest.mark.parametrize("levels", [False, True])
def test_batch_coarse_correction(gpu, levels):
    """Coarse corrections of several instances share one batched launch:
    every solve meets the tolerance, the instance without a coarse space (odd
    n) runs no coarse solve, the pressure solves of the others take fewer
    iterations than the same batch without coarse spaces (six steps from
    rest on small meshes: a few per cent), and the fields
    match the same simulations with the same coarse spaces stepped one by one
    (FP64 at 1e-11, to 1e-7 of the field scale).  With `levels` the 64^2
    and 96^2 instances also get a multigrid level below their coarse space
    and solve their coarse systems by V-cycles, while the 48^2 one stays in
    the batched coarse CG."""
    from navier_stokes_fem_amd import mesh
    from navier_stokes_fem_amd.simulation import step_batch

    specs = [(64, 1.0, 1e-3), (48, 0.7, 2e-3), (33, 1.5, 1e-3), (96, 1.2, 1e-3)]
    rtol = 1e-11

    def with_coarse(sims):
        for (n, _, _), s in zip(specs, sims):
            if n % 2 == 0:
                s.set_coarse_space(mesh.cavity_prolongation(n, 2))
                if levels and n != 48:
                    s.add_coarse_level(mesh.cavity_prolongation(n // 2, 2))
        return sims

    plain = _batch(gpu, specs, "fp64", rtol)
    batched = with_coarse(_batch(gpu, specs, "fp64", rtol))
    one = with_coarse(_batch(gpu, specs, "fp64", rtol))
    st_p = step_batch(plain, 6)
    st_c = step_batch(batched, 6)
    for s in one:
        s.step(6)
    its_p = sum(st["cg_iterations"][2] for (n, _, _), st in zip(specs, st_p) if n % 2 == 0)
    its_c = sum(st["cg_iterations"][2] for (n, _, _), st in zip(specs, st_c) if n % 2 == 0)
    print(f"pressure iterations of the instances with a coarse space: {its_p} -> {its_c}; "
          f"coarse {[st['coarse_iterations'] for st in st_c]}")
    for (n, _, _), st in zip(specs, st_c):
        assert max(st["max_relative_residual"]) <= rtol * 1.01
        assert (st["coarse_iterations"] > 0) == (n % 2 == 0)
    assert its_c < its_p  # a few % this early; the single-mesh tests measure the gain
    for a, b in zip(one, batched):
        for fa, fb in zip(a.fields(), b.fields()):
            np.testing.assert_allclose(fb, fa, rtol=0, atol=1e-7 * max(1.0, np.abs(fa).max()))
    for s in plain + batched + one:
        s.close()
