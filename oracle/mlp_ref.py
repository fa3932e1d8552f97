"""NumPy restatement of the branching MLP (test infrastructure only).

Forward and backward of the 6 -> 64 -> 64 -> 1 ReLU network with the mean
weighted binary cross-entropy used in training, in float64, and a float32
forward pass that rounds weights and activations to BF16 the way the GPU
inference path does (round to nearest even), for checking the GPU kernels.
"""

import numpy as np


def _norm(m, X):
    x = np.zeros((len(X), 8))
    x[:, :6] = (X[:, :6] - m["mean"][:6]) * m["scale"][:6]
    return x


def forward_backward(m, X, y, pos_weight=1.0):
    """Returns (loss, grads) with grads keyed like the model fields."""
    X = np.asarray(X, np.float64)
    y = np.asarray(y, np.float64)
    W1 = np.asarray(m["w1"], np.float64)
    W2 = np.asarray(m["w2"], np.float64)
    b1, b2, w3 = (np.asarray(m[k], np.float64) for k in ("b1", "b2", "w3"))
    b3 = float(m["b3"])
    B = len(y)
    x = _norm(m, X)
    z1 = x @ W1 + b1
    a1 = np.maximum(z1, 0)
    z2 = a1 @ W2 + b2
    a2 = np.maximum(z2, 0)
    out = a2 @ w3 + b3
    w = np.where(y > 0.5, pos_weight, 1.0)
    sp = np.where(out > 0, out + np.log1p(np.exp(-out)), np.log1p(np.exp(out)))
    loss = float(np.sum(w * (sp - y * out)) / B)
    g = w * (1 / (1 + np.exp(-out)) - y) / B
    dz2 = (g[:, None] * w3[None, :]) * (z2 > 0)
    dz1 = (dz2 @ W2.T) * (z1 > 0)
    grads = {"w1": x.T @ dz1, "b1": dz1.sum(0), "w2": a1.T @ dz2, "b2": dz2.sum(0),
             "w3": a2.T @ g, "b3": g.sum()}
    return loss, grads


def bf16(a):
    """Round float32 values to bfloat16 (nearest even) and back."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def infer_bf16(m, X):
    """Inference as the GPU computes it: BF16 inputs, weights and hidden
    activations, float32 accumulation."""
    x = np.zeros((len(X), 8), np.float32)
    x[:, :6] = (np.asarray(X, np.float32)[:, :6] - np.asarray(m["mean"], np.float32)[:6]) \
        * np.asarray(m["scale"], np.float32)[:6]
    x = bf16(x)
    W1 = bf16(m["w1"]).copy()
    W1[6:, :] = 0
    z1 = x @ W1 + np.asarray(m["b1"], np.float32)
    a1 = bf16(np.maximum(z1, 0))
    z2 = a1 @ bf16(m["w2"]) + np.asarray(m["b2"], np.float32)
    a2 = np.maximum(z2, 0)
    return a2 @ np.asarray(m["w3"], np.float32) + np.float32(m["b3"])
