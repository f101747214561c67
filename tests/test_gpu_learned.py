"""GPU tests of the learned pressure warm start (navier_stokes_fem_amd/csrc/
gnn.hip through the C ABI): forward pass and gradients against the CPU
oracle (oracle/gnn.py, which mirrors the kernels' BF16 rounding points and is
itself pinned by finite differences), determinism, training, and the effect
of the trained network on CG iterations."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(gpu, mesh, seed=0):
    from navier_stokes_fem_amd import learned as L
    from navier_stokes_fem_amd.device import DeviceBuffer
    from oracle import gnn
    from oracle import oracle as O

    asm = O.Assembly(mesh, 0.01, 0.01)
    rs, col, val = asm.csr["pressure_stiffness"]
    area = L.node_areas(mesh)
    og = gnn.Graph(rs, col, val, area, mesh.nodes.astype(np.float64))
    pg = L.PressureGraph(gpu, mesh, rs, col, val)
    rng = np.random.default_rng(seed)
    b = rng.standard_normal(og.n) * area * 3.0
    p = rng.standard_normal(og.n) + 0.5
    return og, pg, b, p, DeviceBuffer.from_array(gpu, b), DeviceBuffer.from_array(gpu, p)


def _mesh(kind):
    from navier_stokes_fem_amd import mesh

    return mesh.cavity(12) if kind == "cavity" else mesh.channel(nx=40, ny=10, seed=4)


@pytest.mark.parametrize("kind", ["cavity", "channel"])
def test_forward_matches_oracle(gpu, kind):
    from navier_stokes_fem_amd import learned as L
    from navier_stokes_fem_amd.device import DeviceBuffer
    from oracle import gnn

    og, pg, b, p, bb, _ = _setup(gpu, _mesh(kind))
    params = gnn.init_params(11)
    rng = np.random.default_rng(3)
    for k in params:  # non-zero biases so every ReLU sees both signs
        params[k] = (params[k] + 0.05 * rng.standard_normal(params[k].shape)).astype(np.float32)
    net = L.WarmStartNet(gpu, seed=0)
    net.set_params(L.WarmStartNet.flatten(params))
    np.testing.assert_array_equal(net.get_params(), L.WarmStartNet.flatten(params))
    out = DeviceBuffer(gpu, 8 * og.n)
    net.predict(pg, bb, out)
    got = out.download()
    want, _ = gnn.forward(params, og, b, bf16_mode=True)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f"{kind}: forward relative error vs oracle {err:.2e}")
    assert err < 5e-3
    net.close()
    pg.close()


@pytest.mark.parametrize("kind", ["cavity", "channel"])
def test_gradients_match_oracle(gpu, kind):
    from navier_stokes_fem_amd import learned as L
    from oracle import gnn

    og, pg, b, p, bb, pb = _setup(gpu, _mesh(kind), seed=1)
    params = gnn.init_params(5)
    rng = np.random.default_rng(6)
    for k in params:
        params[k] = (params[k] + 0.05 * rng.standard_normal(params[k].shape)).astype(np.float32)
    net = L.WarmStartNet(gpu)
    net.set_params(L.WarmStartNet.flatten(params))
    loss, flat = net.gradients(pg, bb, pb)
    want_loss, want, _ = gnn.loss_and_grad(params, og, b, p, bf16_mode=True)
    assert loss == pytest.approx(want_loss, rel=5e-3)
    got = L.WarmStartNet.unflatten(flat)
    for name, _ in L.WarmStartNet.LAYOUT:
        e = np.linalg.norm(got[name] - want[name]) / (np.linalg.norm(want[name]) + 1e-30)
        print(f"{kind}: grad {name} relative error {e:.2e}")
        assert e < 1e-2, name
    # the same input gives bit-identical gradients (fixed reduction order)
    loss2, flat2 = net.gradients(pg, bb, pb)
    assert loss2 == loss
    np.testing.assert_array_equal(flat2, flat)
    net.close()
    pg.close()


def test_adam_step_clips_outlier_gradients(gpu):
    """Two Adam steps against the update restated on the host: an ordinary
    pair (its norm starts the running average, no clipping), then the same
    right-hand side with the target scaled by 1e-3, whose relative loss has
    a gradient ~1e5 times larger and is scaled down to 4 times the running
    norm before the moments see it."""
    from navier_stokes_fem_amd import learned as L
    from navier_stokes_fem_amd.device import DeviceBuffer
    from oracle import gnn

    og, pg, b, p, bb, pb = _setup(gpu, _mesh("channel"), seed=2)
    params = gnn.init_params(7)
    rng = np.random.default_rng(8)
    for k in params:
        params[k] = (params[k] + 0.05 * rng.standard_normal(params[k].shape)).astype(np.float32)
    flat0 = L.WarmStartNet.flatten(params)
    net = L.WarmStartNet(gpu)
    net.set_params(flat0)
    small = DeviceBuffer.from_array(gpu, 1e-3 * p)
    lr = 1e-3
    w = flat0.astype(np.float64)
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    avg = 0.0
    for t, target in ((1, pb), (2, small)):
        _, g = net.gradients(pg, bb, target)  # bit-identical to the step's own
        g = g.astype(np.float64)
        norm = float(np.linalg.norm(g))
        avg = avg if avg > 0 else norm
        f = min(1.0, 4.0 * avg / norm)
        print(f"step {t}: gradient norm {norm:.3e}, running {avg:.3e}, scale {f:.3e}")
        assert (f == 1.0) if t == 1 else (f < 1e-3)
        avg = 0.95 * avg + 0.05 * min(norm, 4.0 * avg)
        m = 0.9 * m + 0.1 * f * g
        v = 0.999 * v + 0.001 * (f * g) ** 2
        w = w - lr * (m / (1 - 0.9 ** t)) / (np.sqrt(v / (1 - 0.999 ** t)) + 1e-8)
        net.train_step(pg, bb, target, lr)
        got = net.get_params()
        err = np.abs(got - w).max()
        print(f"step {t}: max |device - host| weight {err:.2e}")
        assert err < 2e-6
    small.close()
    net.close()
    pg.close()


def test_training_reduces_loss(gpu):
    from navier_stokes_fem_amd import learned as L
    from navier_stokes_fem_amd import mesh

    meshes = [mesh.channel(nx=48, ny=12, seed=s, umax=u) for s, u in ((1, 1.0), (2, 1.6))]
    pairs = L.collect_pairs(gpu, meshes, steps=12, warmup=3, dt=5e-3)
    net = L.WarmStartNet(gpu, seed=1)
    before = L.mean_loss(net, pairs)
    hist = L.train(net, pairs, epochs=6, lr=3e-3, lr_final=3e-4)
    after = L.mean_loss(net, pairs)
    print(f"mean loss before {before:.4f}, after {after:.4f}, history {hist}")
    assert before == pytest.approx(1.0)  # the readout starts at zero
    assert after < 0.9 * before
    net.close()


def test_warm_start_keeps_the_solution(gpu):
    """CG solves to the same tolerance whatever the initial guess."""
    from navier_stokes_fem_amd import learned as L
    from navier_stokes_fem_amd import mesh
    from navier_stokes_fem_amd.simulation import Simulation

    m = mesh.cavity(16)
    net = L.WarmStartNet(gpu, seed=2)
    fields = {}
    for mode in ("previous", "zero", "learned", "learned_or_previous", "learned_correction",
                 "extrapolate", "projection"):
        sim = Simulation(m, 0.01, 0.01, device=gpu, rtol=1e-11)
        sim.set_warm_start(mode, net if mode.startswith("learned") else None)
        st = sim.step(4)
        assert max(st["max_relative_residual"]) <= 1e-11 * 1.01
        fields[mode] = sim.fields()
        sim.close()
    ref = fields["previous"]
    for mode, f in fields.items():
        for a, b in zip(f, ref):
            np.testing.assert_allclose(a, b, rtol=0, atol=1e-8 * max(1.0, np.abs(b).max()))
    net.close()


def test_learned_warm_start_gain(gpu):
    """Train on channel flows (inflow speed and jitter seed varied), then on
    a held-out mesh: the trained networks must beat the zero prediction on
    held-out pairs (loss < 1), and pressure CG iterations are measured for
    every initial guess; the learned correction of the previous pressure
    (residual-guarded) must not cost iterations against the reference's
    previous-pressure guess."""
    from navier_stokes_fem_amd import learned as L
    from navier_stokes_fem_amd import mesh

    rng = np.random.default_rng(0)
    train_meshes = [mesh.channel(nx=64, ny=16, seed=100 + s, umax=float(rng.uniform(0.5, 2.0)))
                    for s in range(4)]
    held = mesh.channel(nx=64, ny=16, seed=999, umax=1.3)
    nets = {}
    for target in ("solution", "correction"):
        pairs = L.collect_pairs(gpu, train_meshes, steps=24, warmup=4, dt=2e-3, target=target)
        test_pairs = L.collect_pairs(gpu, [held], steps=8, warmup=4, dt=2e-3, target=target)
        net = L.WarmStartNet(gpu, seed=3)
        assert L.mean_loss(net, test_pairs) == pytest.approx(1.0)  # zero readout
        L.train(net, pairs, epochs=8, lr=3e-3, lr_final=2e-4)
        held_loss = L.mean_loss(net, test_pairs)
        print(f"{target}: held-out loss {held_loss:.4f} (zero prediction = 1)")
        assert held_loss < 1.0
        nets[target] = net
    res = L.evaluate_warm_start(gpu, held, nets["solution"], steps=10, warmup=4, dt=2e-3,
                                modes=("previous", "zero", "learned", "learned_or_previous",
                                       "learned_correction", "learned_search", "projection",
                                       "learned_projection"),
                                correction_net=nets["correction"])
    for mode, r in res.items():
        print(f"{mode:>20}: {r['pressure_iterations_per_step']:.1f} pressure iterations/step, "
              f"{r['steps_per_s']:.1f} steps/s")
    assert all(r["max_relative_residual"] <= 1e-8 * 1.01 for r in res.values())
    it = {mode: r["pressure_iterations_per_step"] for mode, r in res.items()}
    print(f"gain of the network: learned_search vs previous {1 - it['learned_search'] / it['previous']:+.1%}, "
          f"learned_projection vs projection {1 - it['learned_projection'] / it['projection']:+.1%}")
    assert it["learned_correction"] <= 1.02 * it["previous"]
    # the A-norm-optimal step cannot move the guess away from the solution
    assert it["learned_search"] <= 1.02 * it["previous"]
    assert it["learned_projection"] <= 1.02 * it["projection"]
    for net in nets.values():
        net.close()
