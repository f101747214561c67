"""Result files (out_<step>.json) and the time-step convergence estimate of
navier_stokes_fem_amd.results, CPU only.  The estimate is pinned by
Richardson's identity: for runs w_i = w* + C dt_i^q with dt shrinking
tenfold, the order is q and the error estimate is the coarse run's true
error ||w_0 - w*||."""
import json
import os

import numpy as np
import pytest


def test_export_read_round_trip(tmp_path):
    from navier_stokes_fem_amd import results as R

    rng = np.random.default_rng(0)
    u, v, p = rng.standard_normal(37), rng.standard_normal(37), rng.standard_normal(11)
    u[3] = 1e-300
    v[5] = -0.0
    path = R.export_solution(str(tmp_path), 7, u, v, p)
    assert os.path.basename(path) == "out_7.json"
    sol = R.read_solution(path)
    assert sol["timeStep"] == 7
    np.testing.assert_array_equal(sol["u"], u)  # JSON of doubles round-trips exactly
    np.testing.assert_array_equal(sol["v"], v)
    np.testing.assert_array_equal(sol["p"], p)
    with open(path) as f:
        doc = json.load(f)
    assert set(doc) == {"u", "v", "p", "timeStep"}


def test_export_without_folder_writes_nothing():
    from navier_stokes_fem_amd import results as R

    assert R.export_solution("", 3, [1.0], [2.0], [3.0]) is None


def test_read_reference_layout(tmp_path):
    """A file in the reference's layout (keys in the order its JSON library
    writes them, no pressure) reads back."""
    from navier_stokes_fem_amd import results as R

    path = tmp_path / "out_0.json"
    path.write_text('{\n    "timeStep": 0,\n    "u": [\n        0.5,\n        1e-3\n    ],\n'
                    '    "v": [\n        -2.0,\n        0\n    ]\n}\n')
    sol = R.read_solution(str(path))
    np.testing.assert_array_equal(sol["u"], [0.5, 1e-3])
    np.testing.assert_array_equal(sol["v"], [-2.0, 0.0])
    assert sol["p"].size == 0 and sol["timeStep"] == 0
    bad = tmp_path / "bad.json"
    bad.write_text('{"u": [1, 2], "v": [1]}')
    with pytest.raises(ValueError):
        R.read_solution(str(bad))


@pytest.mark.parametrize("q", [1.0, 2.0, 0.5])
def test_runge_estimate_recovers_order_and_error(tmp_path, q):
    from navier_stokes_fem_amd import results as R

    rng = np.random.default_rng(int(10 * q))
    n, steps = 23, 4
    dts = [0.01, 0.001, 0.0001]
    folders = [str(tmp_path / f"run_{i}") for i in range(3)]
    exact, err0 = [], []
    for f in folders:
        os.makedirs(f)
    for k in range(steps):
        ws = rng.standard_normal(2 * n)          # the exact field at this time
        c = rng.standard_normal(2 * n) * (k + 1)  # the leading error term
        for i, f in enumerate(folders):
            w = ws + c * dts[i] ** q
            R.export_solution(f, k * 10 ** i, w[:n], w[n:], np.zeros(3))
        exact.append(ws)
        err0.append(np.linalg.norm(c * dts[0] ** q))
    order, error = R.runge_estimates(folders, steps)
    np.testing.assert_allclose(order, q, rtol=1e-9)
    np.testing.assert_allclose(error, err0, rtol=1e-7)


def test_runge_estimate_degenerate_steps(tmp_path):
    """Equal coarse runs are skipped; b = 0 with a > 0 gives an infinite
    order, as the reference's test of a alone does."""
    from navier_stokes_fem_amd import results as R

    folders = [str(tmp_path / f"r{i}") for i in range(3)]
    for f in folders:
        os.makedirs(f)
    w = np.ones(4)
    for i, f in enumerate(folders):  # step 0: all equal
        R.export_solution(f, 0, w[:2], w[2:], [])
    for i, f in enumerate(folders):  # step 1: run 0 differs, runs 1 and 2 equal
        x = w + (1.0 if i == 0 else 0.0)
        R.export_solution(f, 1 * 10 ** i, x[:2], x[2:], [])
    order, error = R.runge_estimates(folders, 2)
    assert np.isnan(order[0]) and np.isnan(error[0])
    assert np.isinf(order[1]) and np.isnan(error[1])
