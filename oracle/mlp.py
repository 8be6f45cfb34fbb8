"""FP64 warm-start network (test infrastructure; see oracle/__init__.py).

Restates the network the engine's warm start trains (no reference counterpart
in this form: the reference's learned warm starts are torch research modules,
src/temporal_lp_model.py:58-129 and src/supervised_graph_lp.py:13-91, with no
MLP of this shape).  Layers l = 1..L:

    Z_l = H_{l-1} W_l + b_l,   H_l = gelu(Z_l) for l < L,   O = Z_L

with the tanh form gelu(z) = z/2 (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))).
Loss over a batch of M rows and the output split [nx | ny | na]:

    wx/(M nx) sum (O_x - T_x)^2 + wy/(M ny) sum (O_y - T_y)^2
      + wa/(M na) sum softplus(O_a) - T_a O_a

Adam (Kingma & Ba 2015) with bias correction updates the weights.

Pinned by: central finite differences of ``loss`` (the backward matches the
forward it differentiates) and by torch CPU autograd on the same network
(tests/test_oracle_mlp.py); the GEMM kernels by exact integer products.
"""

from __future__ import annotations

import numpy as np

SQRT_2_OVER_PI = np.sqrt(2.0 / np.pi)
GELU_C = 0.044715


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(SQRT_2_OVER_PI * (z + GELU_C * z ** 3)))


def gelu_grad(z):
    t = np.tanh(SQRT_2_OVER_PI * (z + GELU_C * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * SQRT_2_OVER_PI * (1.0 + 3.0 * GELU_C * z * z)


def softplus(o):
    return np.where(o > 0, o + np.log1p(np.exp(-np.abs(o))), np.log1p(np.exp(-np.abs(o))))


def forward(weights, biases, x):
    """Returns the output and the per-layer pre-activations Z and inputs H."""
    h = np.asarray(x, dtype=np.float64)
    zs, hs = [], [h]
    for l, (w, b) in enumerate(zip(weights, biases)):
        z = h @ w + b
        zs.append(z)
        h = gelu(z) if l < len(weights) - 1 else z
        hs.append(h)
    return h, zs, hs


def loss(o, t, nx, ny, na, wx=1.0, wy=1.0, wa=1.0):
    """Mean loss and its gradient with respect to the output."""
    m = o.shape[0]
    g = np.zeros_like(o)
    total = 0.0
    if nx:
        d = o[:, :nx] - t[:, :nx]
        total += wx / (m * nx) * np.sum(d * d)
        g[:, :nx] = 2.0 * wx / (m * nx) * d
    if ny:
        d = o[:, nx:nx + ny] - t[:, nx:nx + ny]
        total += wy / (m * ny) * np.sum(d * d)
        g[:, nx:nx + ny] = 2.0 * wy / (m * ny) * d
    if na:
        oa, ta = o[:, nx + ny:], t[:, nx + ny:]
        total += wa / (m * na) * np.sum(softplus(oa) - ta * oa)
        g[:, nx + ny:] = wa / (m * na) * (1.0 / (1.0 + np.exp(-oa)) - ta)
    return total, g


def backward(weights, zs, hs, d_out):
    """Gradients (dW_l, db_l) of the loss given dL/dO."""
    dws, dbs = [None] * len(weights), [None] * len(weights)
    dz = d_out
    for l in range(len(weights) - 1, -1, -1):
        dws[l] = hs[l].T @ dz
        dbs[l] = dz.sum(axis=0)
        if l > 0:
            dz = (dz @ weights[l].T) * gelu_grad(zs[l - 1])
    return dws, dbs


def loss_and_grad(weights, biases, x, t, nx, ny, na, wx=1.0, wy=1.0, wa=1.0):
    o, zs, hs = forward(weights, biases, x)
    val, g = loss(o, t, nx, ny, na, wx, wy, wa)
    dws, dbs = backward(weights, zs, hs, g)
    return val, dws, dbs


def adam(w, g, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step; returns (w, m, v)."""
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    mh = m / (1.0 - beta1 ** t)
    vh = v / (1.0 - beta2 ** t)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


def bf16_round(a):
    """Round FP32 values to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(np.isnan(f), np.nan, out)
