"""CPU oracle of the learned pressure warm start (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ and by nothing in the product.
The reference has no learned component; this network is specified by this
project (a 3-layer message-passing net over the pressure-matrix graph that
maps the divergence right-hand side to an initial guess for the pressure
CG), and this module is its executable specification:

  graph     A (CSR, the pressure matrix), edge weights
            w_ij = -A_ij / sqrt(A_ii A_jj) for j != i (symmetric), node
            areas a_i (lumped P1 mass), coordinates scaled to the unit box
  input     d_i = b_i / a_i, s = sqrt(mean d^2), x0_i = [d_i / s, X_i, Y_i, 0]
  layer l   m = W_hat h (CSR gather), z = [h | m] W_l + c_l, h' = relu(z)
            (W_1: 8 x 32, W_2, W_3: 64 x 32)
  readout   y_i = h3_i . w4 + c4,   p_hat_i = s * y_i
  loss      L = sum_i (p_hat_i - p_i)^2 / sum_i p_i^2

`bf16=True` rounds exactly where the GPU kernels round (weights, the MFMA
operands [h | m] and dz, stored activations) with round-to-nearest-even, so
GPU and oracle differ only by FP32 summation order; `bf16=False` runs the
same network in FP64, which the finite-difference test pins.
"""
from __future__ import annotations

import numpy as np

HIDDEN = 32
IN_FEATURES = 4  # d/s, X, Y, 0 (padding)
LAYER_IN = (2 * IN_FEATURES, 2 * HIDDEN, 2 * HIDDEN)


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), a, out).astype(np.float32)


class Graph:
    """Pressure-matrix graph of one mesh (CSR without the diagonal)."""

    def __init__(self, row_start, col, val, area, xy):
        rs = np.asarray(row_start, np.int64)
        col = np.asarray(col, np.int64)
        val = np.asarray(val, np.float64)
        n = len(rs) - 1
        rows = np.repeat(np.arange(n), np.diff(rs))
        diag = np.zeros(n)
        on = col == rows
        diag[rows[on]] = val[on]
        off = ~on
        self.n = n
        self.rows = rows[off]
        self.cols = col[off]
        self.w = -val[off] / np.sqrt(diag[self.rows] * diag[self.cols])
        self.area = np.asarray(area, np.float64)
        xy = np.asarray(xy, np.float64)[:n]
        lo = xy.min(0)
        ext = max(float((xy.max(0) - lo).max()), 1e-30)
        self.X = (xy - lo) / ext

    def agg(self, h):
        out = np.zeros_like(h)
        np.add.at(out, self.rows, self.w[:, None] * h[self.cols])
        return out


def init_params(seed=0):
    """He-style initialisation (FP32 master weights)."""
    rng = np.random.default_rng(seed)
    p = {}
    for l, k in enumerate(LAYER_IN, start=1):
        p[f"W{l}"] = (rng.standard_normal((k, HIDDEN)) * np.sqrt(2.0 / k)).astype(np.float32)
        p[f"c{l}"] = np.zeros(HIDDEN, np.float32)
    p["w4"] = (rng.standard_normal(HIDDEN) * np.sqrt(1.0 / HIDDEN)).astype(np.float32)
    p["c4"] = np.zeros(1, np.float32)
    return p


def _features(g: Graph, b):
    d = np.asarray(b, np.float64) / g.area
    s = float(np.sqrt(np.mean(d * d))) + 1e-30
    x0 = np.zeros((g.n, IN_FEATURES))
    x0[:, 0] = d / s
    x0[:, 1:3] = g.X
    return x0, s


def forward(params, g: Graph, b, bf16_mode=True):
    """Returns p_hat and the cache the backward pass needs."""
    dt = np.float32 if bf16_mode else np.float64
    R = bf16 if bf16_mode else (lambda a: np.asarray(a, np.float64))
    x0, s = _features(g, b)
    h = R(x0.astype(dt))
    cache = {"s": s, "layers": []}
    for l in (1, 2, 3):
        W = R(params[f"W{l}"])
        m = g.agg(h.astype(np.float64)).astype(dt)
        X = R(np.concatenate([h, m], axis=1))
        z = (X.astype(np.float64) @ W.astype(np.float64)).astype(dt) + params[f"c{l}"].astype(dt)
        hn = np.maximum(z, 0)
        cache["layers"].append((X, W, z))
        h = R(hn)
    y = (h.astype(np.float64) @ params["w4"].astype(np.float64) + float(params["c4"][0])).astype(dt)
    cache["h3"] = h
    return s * y.astype(np.float64), cache


def loss_and_grad(params, g: Graph, b, p, bf16_mode=True):
    """Loss and gradients of every parameter (same names as params)."""
    dt = np.float32 if bf16_mode else np.float64
    R = bf16 if bf16_mode else (lambda a: np.asarray(a, np.float64))
    p = np.asarray(p, np.float64)
    p_hat, cache = forward(params, g, b, bf16_mode)
    s = cache["s"]
    den = float(np.sum(p * p)) + 1e-300
    diff = p_hat - p
    loss = float(np.sum(diff * diff)) / den
    dy = (2.0 * s * diff / den).astype(dt)
    grads = {}
    h3 = cache["h3"]
    grads["w4"] = (h3.astype(np.float64).T @ dy.astype(np.float64)).astype(np.float32)
    grads["c4"] = np.array([np.sum(dy.astype(np.float64))], np.float32)
    dh = (dy[:, None].astype(np.float64) * params["w4"][None, :].astype(np.float64)).astype(dt)
    for l in (3, 2, 1):
        X, W, z = cache["layers"][l - 1]
        dz = R(np.where(z > 0, dh, 0).astype(dt))
        grads[f"W{l}"] = (X.astype(np.float64).T @ dz.astype(np.float64)).astype(np.float32)
        grads[f"c{l}"] = np.sum(dz.astype(np.float64), axis=0).astype(np.float32)
        if l == 1:
            break
        dX = (dz.astype(np.float64) @ W.astype(np.float64).T).astype(dt)
        k = dX.shape[1] // 2
        dh = (dX[:, :k].astype(np.float64) + g.agg(dX[:, k:].astype(np.float64))).astype(dt)
    return loss, grads, p_hat
