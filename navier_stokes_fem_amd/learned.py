"""Learned warm start of the pressure solve (no counterpart in the reference,
whose pressure CG starts from the previous pressure, include/assembly.h:
608-615).

A 3-layer message-passing network over the pressure-matrix graph maps the
divergence right-hand side b to an initial guess for the pressure CG; the CG
then runs to the same tolerance, so the guess can only change the iteration
count.  Forward, backward and Adam run in libnsfem_hip.so (``nsfem_gnn_*``,
BF16 MFMA dense layers with FP32 accumulation); this module builds the
graphs from meshes, collects (b, p) training pairs from simulations that the
plain solver runs, trains, and measures the effect on CG iterations.

    net = WarmStartNet(device, seed=0)
    pairs = collect_pairs(device, [mesh.channel(512, 128, seed=s, umax=u) ...])
    train(net, pairs, epochs=4)
    sim.set_warm_start("learned_or_previous", net)
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _native as N
from .device import DeviceBuffer, assemble

__all__ = ["PressureGraph", "WarmStartNet", "node_areas", "collect_pair_sets", "collect_pairs",
           "mean_loss", "train", "evaluate_warm_start"]


def node_areas(mesh) -> np.ndarray:
    """Lumped P1 mass of each pressure node: a third of the area of every
    triangle the node is a vertex of."""
    xy = mesh.nodes.astype(np.float64)
    t = mesh.elements[:, :3]
    a, b, c = xy[t[:, 0]], xy[t[:, 1]], xy[t[:, 2]]
    area = 0.5 * np.abs((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) -
                        (c[:, 0] - a[:, 0]) * (b[:, 1] - a[:, 1]))
    out = np.zeros(mesh.pressure_count)
    for k in range(3):
        np.add.at(out, t[:, k], area / 3.0)
    return out


class PressureGraph:
    """The network's view of one mesh: its pressure matrix as a graph."""

    def __init__(self, device, mesh, row_start, col, val):
        self.device = device
        self.n = mesh.pressure_count
        rs = np.ascontiguousarray(row_start, np.int32)
        cl = np.ascontiguousarray(col, np.int32)
        vl = np.ascontiguousarray(val, np.float64)
        area = node_areas(mesh)
        xy = np.ascontiguousarray(mesh.nodes[:self.n], np.float32).reshape(-1)
        h = C.c_void_p()
        N.check(N.lib().nsfem_gnn_graph_create(device.handle, self.n, len(cl), N.ptr(rs), N.ptr(cl),
                                               N.ptr(vl), N.ptr(area), N.ptr(xy), C.byref(h)))
        self.handle = h

    @classmethod
    def from_mesh(cls, device, mesh, viscosity=1e-3, dt=1e-3):
        rs, col, val, _ = assemble(mesh, viscosity, dt)["pressure_stiffness"]
        return cls(device, mesh, rs, col, val)

    def close(self):
        if getattr(self, "handle", None):
            N.lib().nsfem_gnn_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class WarmStartNet:
    """The network (parameters and Adam state live on the device)."""

    PARAM_COUNT = N.GNN_PARAM_COUNT
    LAYOUT = (("W1", (8, 32)), ("c1", (32,)), ("W2", (64, 32)), ("c2", (32,)),
              ("W3", (64, 32)), ("c3", (32,)), ("w4", (32,)), ("c4", (1,)))

    def __init__(self, device, seed: int = 0):
        self.device = device
        h = C.c_void_p()
        N.check(N.lib().nsfem_gnn_create(device.handle, int(seed), C.byref(h)))
        self.handle = h

    # -- parameters -----------------------------------------------------
    def get_params(self) -> np.ndarray:
        out = np.empty(self.PARAM_COUNT, np.float32)
        N.check(N.lib().nsfem_gnn_get_params(self.handle, N.ptr(out)))
        return out

    def set_params(self, flat) -> None:
        flat = np.ascontiguousarray(flat, np.float32)
        if flat.size != self.PARAM_COUNT:
            raise ValueError(f"expected {self.PARAM_COUNT} parameters, got {flat.size}")
        N.check(N.lib().nsfem_gnn_set_params(self.handle, N.ptr(flat)))

    @classmethod
    def unflatten(cls, flat) -> dict:
        out, k = {}, 0
        for name, shape in cls.LAYOUT:
            size = int(np.prod(shape))
            out[name] = np.asarray(flat[k:k + size], np.float32).reshape(shape)
            k += size
        return out

    @classmethod
    def flatten(cls, params: dict) -> np.ndarray:
        return np.concatenate([np.asarray(params[name], np.float32).reshape(-1)
                               for name, _ in cls.LAYOUT])

    def save(self, path):
        np.savez(path, params=self.get_params())

    def load(self, path):
        with np.load(path, allow_pickle=False) as z:
            self.set_params(z["params"])

    # -- compute ----------------------------------------------------------
    def predict(self, graph: PressureGraph, b: DeviceBuffer, out: DeviceBuffer) -> None:
        N.check(N.lib().nsfem_gnn_predict(self.handle, graph.handle, b.handle, out.handle))

    def gradients(self, graph: PressureGraph, b: DeviceBuffer, p: DeviceBuffer):
        g = np.empty(self.PARAM_COUNT, np.float32)
        loss = C.c_double(0.0)
        N.check(N.lib().nsfem_gnn_gradients(self.handle, graph.handle, b.handle, p.handle,
                                            N.ptr(g), C.byref(loss)))
        return loss.value, g

    def train_step(self, graph: PressureGraph, b: DeviceBuffer, p: DeviceBuffer, lr: float,
                   sync: bool = False):
        loss = C.c_double(0.0)
        N.check(N.lib().nsfem_gnn_train_step(self.handle, graph.handle, b.handle, p.handle,
                                             float(lr), C.byref(loss) if sync else None))
        return loss.value if sync else None

    def close(self):
        if getattr(self, "handle", None):
            N.lib().nsfem_gnn_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def collect_pair_sets(device, meshes, viscosity=1e-3, dt=5e-4, steps=40, warmup=5, every=1,
                      precision="fp64", targets=("solution",), log=None, log_every=50):
    """Run the plain solver on every mesh and keep, per mesh, one training
    pair per pressure solve of steps warmup+1 .. warmup+steps (every
    `every`-th), for each requested target:

    * "solution":   (graph, b, p): the right-hand side and the pressure CG's
      solution (for the ``learned`` modes);
    * "correction": (graph, r, delta) with r = b - A p_prev and
      delta = p - p_prev, p_prev the previous step's pressure (for the
      ``learned_correction`` mode).

    All meshes step together (``step_batch``: each solve kind of every mesh
    in one persistent launch), so the pairs of many meshes cost little more
    than those of one.  Returns {target: [pairs]}; pairs of one mesh are
    contiguous, meshes in the given order."""
    from .simulation import Simulation, step_batch

    targets = tuple(targets)
    if not targets or any(t not in ("solution", "correction") for t in targets):
        raise ValueError("targets must be a non-empty subset of ('solution', 'correction')")
    sims, graphs, prevs = [], [], []
    per_mesh = {t: [[] for _ in meshes] for t in targets}
    try:
        for m in meshes:
            sims.append(Simulation(m, viscosity, dt, device=device, precision=precision))
            graphs.append(PressureGraph.from_mesh(device, m, viscosity, dt))
            if "correction" in targets:
                prevs.append(DeviceBuffer(device, 8 * m.pressure_count))

        def advance():
            if len(sims) > 1:
                step_batch(sims, 1)
            else:
                sims[0].step(1)

        for _ in range(warmup):
            advance()
        t0 = time.perf_counter()
        for k in range(steps):
            for s, prev in zip(sims, prevs):
                s.copy_pressure_system(None, prev)
            advance()
            if k % every:
                continue
            for i, (m, s, g) in enumerate(zip(meshes, sims, graphs)):
                P = m.pressure_count
                b = DeviceBuffer(device, 8 * P)
                p = DeviceBuffer(device, 8 * P)
                s.copy_pressure_system(b, p)
                if "correction" in targets:
                    r = DeviceBuffer(device, 8 * P)
                    d = DeviceBuffer(device, 8 * P)
                    s.sp_r_mult(0, prevs[i], r, lhs=b)               # r <- b - A p_prev
                    device.saxpby(P, 1.0, -1.0, p, prevs[i], d)      # d <- p - p_prev
                    per_mesh["correction"][i].append((g, r, d))
                if "solution" in targets:
                    per_mesh["solution"][i].append((g, b, p))
                else:
                    b.close()
                    p.close()
            if log and (k + 1) % log_every == 0:
                log(f"collected step {k + 1}/{steps} of {len(sims)} meshes "
                    f"({time.perf_counter() - t0:.1f} s)")
        device.synchronize()
    finally:
        for prev in prevs:
            prev.close()
        for s in sims:
            s.close()
    return {t: [pair for mesh_pairs in per_mesh[t] for pair in mesh_pairs] for t in targets}


def collect_pairs(device, meshes, viscosity=1e-3, dt=5e-4, steps=40, warmup=5, every=1,
                  precision="fp64", target="solution", log=None):
    """``collect_pair_sets`` for one target: the list of its pairs."""
    if target not in ("solution", "correction"):
        raise ValueError("target must be 'solution' or 'correction'")
    return collect_pair_sets(device, meshes, viscosity, dt, steps, warmup, every, precision,
                             (target,), log)[target]


def mean_loss(net, pairs):
    """Mean of the loss sum (p_hat - p)^2 / sum p^2 over pairs (1 = the zero
    prediction)."""
    return float(np.mean([net.gradients(g, b, p)[0] for g, b, p in pairs]))


def train(net: WarmStartNet, pairs, epochs=4, lr=2e-3, lr_final=2e-4, seed=0, log=None,
          log_every=500):
    """Adam over shuffled pairs with a cosine learning-rate decay; returns
    the per-epoch mean loss (measured on every pair after the epoch)."""
    rng = np.random.default_rng(seed)
    total = epochs * len(pairs)
    step = 0
    history = []
    t0 = time.perf_counter()
    for ep in range(epochs):
        order = rng.permutation(len(pairs))
        for i in order:
            g, b, p = pairs[i]
            frac = step / max(1, total - 1)
            rate = lr_final + 0.5 * (lr - lr_final) * (1.0 + np.cos(np.pi * frac))
            sync = log is not None and step % log_every == 0
            loss = net.train_step(g, b, p, rate, sync=sync)
            if sync:
                log(f"epoch {ep + 1}/{epochs} step {step}/{total}: loss {loss:.4f} "
                    f"lr {rate:.2e} ({time.perf_counter() - t0:.1f} s)")
            step += 1
        history.append(mean_loss(net, pairs[::max(1, len(pairs) // 64)]))
        if log:
            log(f"epoch {ep + 1}: mean loss {history[-1]:.4f}")
    return history


def evaluate_warm_start(device, mesh, net, modes=("previous", "zero", "learned",
                                                  "learned_or_previous"), correction_net=None,
                        viscosity=1e-3, dt=5e-4, steps=20, warmup=5, precision="fp64"):
    """Same simulation, same steps, one run per initial-guess mode: pressure
    CG iterations per step, pressure CG seconds per step and steps/s."""
    from .simulation import Simulation

    sim = Simulation(mesh, viscosity, dt, device=device, precision=precision)
    out = {}
    for mode in modes:
        sim.reset()
        sim.set_warm_start("previous")
        if warmup:
            sim.step(warmup)
        model = (correction_net if mode in ("learned_correction", "learned_search",
                                             "learned_projection") else net)
        sim.set_warm_start(mode, model if mode.startswith("learned") else None)  # noqa
        device.synchronize()
        t = time.perf_counter()
        st = sim.step(steps)
        wall = time.perf_counter() - t
        out[mode] = {"pressure_iterations_per_step": st["cg_iterations"][2] / steps,
                     "pressure_seconds_per_step": st["cg_seconds"][2] / steps,
                     "steps_per_s": steps / wall,
                     "max_relative_residual": st["max_relative_residual"][2]}
    sim.set_warm_start("previous")
    sim.close()
    return out
