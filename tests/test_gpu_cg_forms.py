"""Forms of the one-column mixed CG iteration against each other: one load
per nonzero per lane (items) and eight consecutive nonzeros per lane
(groups).  They compute the same products and sums in the same order, so solutions, iteration counts and
residuals are bit-identical; the 768-thread block variant partitions the
rows differently (other reduction order) and agrees to the solve's
tolerance.  The form is chosen by NSFEM_CG_PHASE /
NSFEM_CG_BLOCK, which the library reads at every solve; without them a
single solve runs group loads in 768-thread blocks."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def SYSTEMS():
    from navier_stokes_fem_amd import mesh

    return (("cavity96", mesh.cavity(96), "pressure_stiffness"),
            ("channel64", mesh.channel(64, 16, seed=3), "diffusion"),
            ("cavity40", mesh.cavity(40), "velocity_mass"))


def _solves(gpu, phase, block="1024", grid=None, systems=None):
    from navier_stokes_fem_amd.device import assemble

    old = {k: os.environ.get(k) for k in ("NSFEM_CG_PHASE", "NSFEM_CG_BLOCK", "NSFEM_CG_GRID")}
    for k, v in (("NSFEM_CG_PHASE", phase), ("NSFEM_CG_BLOCK", block), ("NSFEM_CG_GRID", grid)):
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v
    out = {}
    try:
        for name, m, key in systems or SYSTEMS():
            rs, col, val, shape = assemble(m, 1e-3, 1e-3)[key]
            gpu.upload_csr(0, rs, col, val, shape=shape)
            n = shape[0]
            b = gpu.buffer(np.random.default_rng(3).standard_normal(n))
            x = gpu.buffer(np.zeros(n))
            st = gpu.cg_solve(0, b, x, rtol=1e-10, precision="fp32_refined")[0]
            out[name] = (x.download(), st["iterations"], st["refinements"], st["final_residual"],
                         st["rhs_norm"])
            b.close()
            x.close()
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return out


@pytest.mark.parametrize("block", ["1024", "768"])
def test_phase_forms_are_bit_identical(gpu, block):
    ref = _solves(gpu, "items", block)
    for phase in ("groups",):
        got = _solves(gpu, phase, block)
        for k in ref:
            np.testing.assert_array_equal(got[k][0], ref[k][0], err_msg=f"{phase} {block} {k}")
            assert got[k][1:] == ref[k][1:], (phase, block, k, got[k][1:], ref[k][1:])
    print({k: v[1:3] for k, v in ref.items()})


def test_phase_forms_with_descriptors_beyond_lds(gpu):
    """One block over a P2 matrix of ~6,000 wave tiles: more tile
    descriptors than the block's LDS copy holds, so every form reads them
    from global memory."""
    from navier_stokes_fem_amd import mesh

    systems = (("cavity200_diffusion", mesh.cavity(200), "diffusion"),)
    ref = _solves(gpu, "items", "768", grid="1", systems=systems)
    for phase in ("groups",):
        got = _solves(gpu, phase, "768", grid="1", systems=systems)
        for k in ref:
            np.testing.assert_array_equal(got[k][0], ref[k][0], err_msg=f"{phase} {k}")
            assert got[k][1:] == ref[k][1:], (phase, k, got[k][1:], ref[k][1:])


def test_default_form_is_groups_in_wide_blocks(gpu):
    ref = _solves(gpu, "groups", "768")
    got = _solves(gpu, None, None)
    for k in ref:
        np.testing.assert_array_equal(got[k][0], ref[k][0], err_msg=k)
        assert got[k][1:] == ref[k][1:], (k, got[k][1:], ref[k][1:])


@pytest.mark.parametrize("phase", ["items", "groups"])
def test_wide_blocks_solve_to_the_tolerance(gpu, phase):
    ref = _solves(gpu, phase)
    got = _solves(gpu, phase, block="768")
    for k in ref:
        x, it, _, res, bn = got[k]
        assert res <= 1e-10 * bn * (1 + 1e-9), (k, res, bn)
        scale = np.abs(ref[k][0]).max()
        assert np.abs(x - ref[k][0]).max() <= 1e-5 * scale, k
        assert abs(it - ref[k][1]) <= max(3, 0.05 * ref[k][1]), (k, it, ref[k][1])
