"""Pins of the oracle's composed time step (or_step, the restatement of
include/assembly.h:1516-1554 with diffusionSolve :489-558 and applyPressure
:560-712).

The parts of or_step are pinned one by one elsewhere (CG, products,
assembly, advection); these tests pin their composition -- the signs and dt
scalings of the divergence and gradient, the Dirichlet right-hand-side
patches, the velocity update and the advection's place in the step -- with
flows whose outcome the mathematics fixes (tests/step_cases.py).  Each check
is also run against every planted mistake of the oracle
(oracle.STEP_FAULTS, or_set_step_fault): the faithful step passes all of
them and every mistake fails at least one.
"""
import numpy as np
import pytest

import step_cases as S


def oracle_stepper(O, rtol=1e-10):
    cache = {}

    def stepper(m, nu, dt, u, v, p, steps):
        key = (id(m), nu, dt)
        if key not in cache:
            cache.clear()
            cache[key] = (m, O.Assembly(m, nu, dt), O.kd_build(m.xy(), m.elems()))
        _, asm, tree = cache[key]
        for _ in range(steps):
            u, v, p, st = O.step(asm, tree, m, dt, u, v, p, rtol=rtol)
            assert st["status"] == 0
        return u, v, p

    return stepper


@pytest.mark.parametrize("check", [S.check_uniform_flow, S.check_potential_flow,
                                   S.check_rotation, S.check_runge_order,
                                   S.check_mirror_symmetry, S.check_dirichlet_values],
                         ids=lambda f: f.__name__[6:])
def test_faithful_step_passes(oracle, check):
    ok, detail = check(oracle_stepper(oracle))
    assert ok, detail


@pytest.mark.parametrize("fault", list(range(1, 10)),
                         ids=lambda k: ("div_sign", "grad_sign", "vel_bc_sign", "p_bc_sign",
                                        "div_scale", "grad_scale", "no_update_add",
                                        "skip_advection", "swap_components")[k - 1])
def test_every_planted_mistake_is_caught(oracle, fault):
    name = oracle.STEP_FAULTS[fault - 1]
    failed = []
    with oracle.step_fault(name):
        assert oracle.lib().or_get_step_fault() == fault
        for check in S.COMPOSITION_CHECKS:
            ok, detail = check(oracle_stepper(oracle))
            if not ok:
                failed.append(detail)
    assert oracle.lib().or_get_step_fault() == 0
    assert failed, f"no composition check caught the planted mistake {name}"


def test_mirror_symmetry_catches_swapped_components(oracle):
    with oracle.step_fault("swap_components"):
        ok, detail = S.check_mirror_symmetry(oracle_stepper(oracle))
    assert not ok, detail


def test_potential_flow_error_converges(oracle):
    """The faithful step's projection error is a discretisation error: it
    falls under mesh refinement (16 -> 32 -> 64: p 3.2% -> 0.92% -> 0.25%,
    interior u 6.4% -> 2.6% -> 1.3%)."""
    st = oracle_stepper(oracle)
    e = [S.potential_flow_errors(st, n) for n in (16, 32, 64)]
    assert e[1][0] < 0.5 * e[0][0] and e[2][0] < 0.5 * e[1][0], e
    assert e[1][1] < 0.6 * e[0][1] and e[2][1] < 0.6 * e[1][1], e
