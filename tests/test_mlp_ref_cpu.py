"""Pins of the NumPy MLP reference that the GPU kernels are checked
against: its gradients against central finite differences of its own loss,
its loss against torch's BCE-with-logits, its BF16 rounding against torch's
float32 -> bfloat16 conversion, and its BF16 inference against the float64
forward within BF16 error."""

import numpy as np
import pytest

from oracle.mlp_ref import bf16, forward_backward, infer_bf16


def model(seed):
    rng = np.random.default_rng(seed)
    m = {"w1": rng.normal(0, 0.4, (8, 64)), "b1": rng.normal(0, 0.1, 64),
         "w2": rng.normal(0, 0.15, (64, 64)), "b2": rng.normal(0, 0.1, 64),
         "w3": rng.normal(0, 0.2, 64), "b3": 0.1,
         "mean": np.array([5, 50, 2, 20, 10, 15, 0, 0], float),
         "scale": np.array([0.2, 0.02, 0.5, 0.05, 0.1, 0.08, 0, 0], float)}
    return m


def data(n, seed):
    rng = np.random.default_rng(seed)
    X = np.stack([rng.integers(0, 20, n), rng.integers(0, 200, n), rng.integers(1, 6, n),
                  rng.integers(1, 60, n), rng.integers(1, 30, n), rng.random(n) * 4], 1).astype(float)
    y = (rng.random(n) < 0.3).astype(float)
    return X, y


def test_gradients_equal_finite_differences():
    m = model(0)
    X, y = data(64, 1)
    _, g = forward_backward(m, X, y, pos_weight=2.5)
    rng = np.random.default_rng(2)
    eps = 1e-6
    for k in ("w1", "b1", "w2", "b2", "w3", "b3"):
        for _ in range(6):
            mp = {kk: np.array(v, float, copy=True) for kk, v in m.items()}
            mm = {kk: np.array(v, float, copy=True) for kk, v in m.items()}
            if k == "b3":
                idx = ()
            else:
                shape = np.shape(m[k])
                idx = tuple(int(rng.integers(0, s)) for s in shape)
                if k == "w1":
                    idx = (int(rng.integers(0, 6)), idx[1])  # rows 6,7 are padding
            mp[k][idx] += eps
            mm[k][idx] -= eps
            lp, _ = forward_backward(mp, X, y, pos_weight=2.5)
            lm, _ = forward_backward(mm, X, y, pos_weight=2.5)
            fd = (lp - lm) / (2 * eps)
            an = float(np.asarray(g[k])[idx])
            assert abs(fd - an) <= 1e-6 + 1e-4 * abs(fd), (k, idx, fd, an)


def test_loss_equals_torch_bce_with_logits():
    torch = pytest.importorskip("torch")
    m = model(3)
    X, y = data(200, 4)
    loss, _ = forward_backward(m, X, y, pos_weight=3.0)
    x = np.zeros((len(X), 8))
    x[:, :6] = (X[:, :6] - m["mean"][:6]) * m["scale"][:6]
    t = lambda a: torch.tensor(np.asarray(a, float), dtype=torch.float64)
    a1 = torch.relu(t(x) @ t(m["w1"]) + t(m["b1"]))
    a2 = torch.relu(a1 @ t(m["w2"]) + t(m["b2"]))
    out = a2 @ t(m["w3"]) + m["b3"]
    w = torch.where(t(y) > 0.5, 3.0, 1.0)
    ref = torch.nn.functional.binary_cross_entropy_with_logits(out, t(y), weight=w)
    assert abs(loss - float(ref)) < 1e-10


def test_bf16_rounding_equals_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    a = np.concatenate([rng.normal(0, 10, 5000), rng.normal(0, 1e-3, 5000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.01171875]]).astype(np.float32)
    ref = torch.tensor(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(bf16(a).view(np.uint32), ref.view(np.uint32))


def test_bf16_inference_close_to_float64_forward():
    m = model(6)
    X, _ = data(500, 7)
    got = infer_bf16(m, X)
    x = np.zeros((len(X), 8))
    x[:, :6] = (X[:, :6] - m["mean"][:6]) * m["scale"][:6]
    a1 = np.maximum(x @ m["w1"] + m["b1"], 0)
    a2 = np.maximum(a1 @ m["w2"] + m["b2"], 0)
    ref = a2 @ m["w3"] + m["b3"]
    # BF16 keeps 8 significant bits: relative error per operand 2^-9
    scale = np.abs(ref).max()
    assert np.max(np.abs(got - ref)) < 0.03 * max(1.0, scale)
    # a plausible mistake (a dropped bias) is far outside that tolerance
    m2 = dict(m, b2=np.zeros(64))
    assert np.max(np.abs(infer_bf16(m2, X) - ref)) > 0.03 * max(1.0, scale)
