"""Pins of the learned warm start's CPU oracle (oracle/gnn.py), CPU only:
the FP64 network's gradients against central finite differences, the BF16
rounding against torch's bfloat16 conversion, and the BF16-mirroring network
(what the GPU kernels compute) against the FP64 one."""
import numpy as np
import pytest


def _graph(n=6, seed=0):
    from navier_stokes_fem_amd import mesh
    from oracle import gnn
    from oracle import oracle as O

    m = mesh.cavity(n)
    asm = O.Assembly(m, 0.01, 0.01)
    rs, col, val = asm.csr["pressure_stiffness"]
    xy = m.nodes.astype(np.float64)
    area = np.zeros(m.pressure_count)
    tri = xy[m.elements[:, :3]]
    a = 0.5 * np.abs((tri[:, 1, 0] - tri[:, 0, 0]) * (tri[:, 2, 1] - tri[:, 0, 1])
                     - (tri[:, 2, 0] - tri[:, 0, 0]) * (tri[:, 1, 1] - tri[:, 0, 1]))
    for k in range(3):
        np.add.at(area, m.elements[:, k], a / 3.0)
    g = gnn.Graph(rs, col, val, area, xy)
    rng = np.random.default_rng(seed)
    b = rng.standard_normal(g.n) * area
    p = rng.standard_normal(g.n)
    return g, b, p


def test_bf16_rounding_matches_torch():
    import torch

    from oracle import gnn

    rng = np.random.default_rng(1)
    a = np.concatenate([rng.standard_normal(5000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 5000),
                        np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -0.0, 65504.0], np.float32)])
    a = a.astype(np.float32)
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(gnn.bf16(a), ref)


def test_fp64_gradients_match_finite_differences():
    from oracle import gnn

    g, b, p = _graph(5)
    params = {k: v.astype(np.float64) for k, v in gnn.init_params(3).items()}
    # move the biases off zero so every ReLU sees both signs
    rng = np.random.default_rng(4)
    for k in params:
        params[k] = params[k] + 0.05 * rng.standard_normal(params[k].shape)
    loss, grads, _ = gnn.loss_and_grad(params, g, b, p, bf16_mode=False)
    assert loss > 0
    eps = 1e-6
    for name, val in params.items():
        flat = val.reshape(-1)
        idx = rng.choice(flat.size, size=min(12, flat.size), replace=False)
        for i in idx:
            old = flat[i]
            flat[i] = old + eps
            lp, _, _ = gnn.loss_and_grad(params, g, b, p, bf16_mode=False)
            flat[i] = old - eps
            lm, _, _ = gnn.loss_and_grad(params, g, b, p, bf16_mode=False)
            flat[i] = old
            fd = (lp - lm) / (2 * eps)
            an = float(grads[name].reshape(-1)[i])
            assert an == pytest.approx(fd, rel=1e-4, abs=1e-7), (name, i, an, fd)


def test_bf16_network_tracks_fp64():
    from oracle import gnn

    g, b, p = _graph(8, seed=2)
    params = gnn.init_params(5)
    l32, g32, y32 = gnn.loss_and_grad(params, g, b, p, bf16_mode=True)
    l64, g64, y64 = gnn.loss_and_grad({k: v.astype(np.float64) for k, v in params.items()},
                                      g, b, p, bf16_mode=False)
    assert l32 == pytest.approx(l64, rel=2e-2)
    assert np.linalg.norm(y32 - y64) <= 2e-2 * np.linalg.norm(y64)
    for k in g64:
        assert np.linalg.norm(g32[k] - g64[k]) <= 5e-2 * np.linalg.norm(g64[k]) + 1e-8, k


def test_prediction_is_scale_equivariant():
    """p_hat(c b) = c p_hat(b): the input is normalised by its own scale."""
    from oracle import gnn

    g, b, _ = _graph(6, seed=3)
    params = {k: v.astype(np.float64) for k, v in gnn.init_params(7).items()}
    y1, _ = gnn.forward(params, g, b, bf16_mode=False)
    y2, _ = gnn.forward(params, g, 1e-3 * b, bf16_mode=False)
    np.testing.assert_allclose(y2, 1e-3 * y1, rtol=1e-9, atol=1e-15)
