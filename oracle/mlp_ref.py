"""CPU reference of the engine's cost-model MLP (numpy).

TEST INFRASTRUCTURE ONLY: the checker for the GPU forward / backward kernels.
Network: 8 -> 32 (ReLU) -> 32 (ReLU) -> 2, weights used in BF16 (the fp32
master copy rounded to nearest-even), activations and sums in the working
precision (float32, or float64 for finite-difference checks).  Parameter
layout: W1[8][32], b1[32], W2[32][32], b2[32], W3[32][2], b3[2].
Loss: sum M * (y - T)^2 / sum M, with T already standardised.

Pinned by finite differences in float64 (tests/test_mlp.py).
"""
import numpy as np

F, H, O = 8, 32, 2
SHAPES = [("W1", (F, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, O)), ("b3", (O,))]
P = sum(int(np.prod(s)) for _, s in SHAPES)


def bf16_round(x):
    """float32 -> bf16 -> float32, round to nearest even (finite values)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(p, dtype=np.float32, bf16=True):
    p = np.asarray(p, np.float32)
    if bf16:
        p = bf16_round(p)
    out, o = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = p[o:o + k].reshape(shape).astype(dtype)
        o += k
    return out


def forward(params, X, norm=None, dtype=np.float32):
    W = unpack(params, dtype)
    X = np.asarray(X, dtype)
    if norm is not None:
        xm, xs, _, _ = norm
        X = (X - np.asarray(xm, dtype)) / np.asarray(xs, dtype)
    a1 = X @ W["W1"] + W["b1"]
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W["W2"] + W["b2"]
    h2 = np.maximum(a2, 0)
    y = h2 @ W["W3"] + W["b3"]
    return y, (X, h1, h2)


def predict(params, X, norm):
    y, _ = forward(params, X, norm)
    _, _, ym, ys = norm
    return y * np.asarray(ys, np.float32) + np.asarray(ym, np.float32)


def loss_grad(params, X, T, M, norm=None, dtype=np.float32):
    """Loss and gradient w.r.t. the parameters (straight through the BF16
    rounding).  T is standardised with norm's y statistics when norm is given."""
    T = np.asarray(T, dtype)
    M = np.asarray(M, dtype)
    if norm is not None:
        _, _, ym, ys = norm
        T = (T - np.asarray(ym, dtype)) / np.asarray(ys, dtype)
    W = unpack(params, dtype)
    y, (x, h1, h2) = forward(params, X, norm, dtype)
    msum = M.sum()
    e = y - T
    loss = (M * e * e).sum() / msum
    dy = M * 2 * e / msum
    g = {"W3": h2.T @ dy, "b3": dy.sum(0)}
    d2 = (dy @ W["W3"].T) * (h2 > 0)
    g["W2"] = h1.T @ d2
    g["b2"] = d2.sum(0)
    d1 = (d2 @ W["W2"].T) * (h1 > 0)
    g["W1"] = x.T @ d1
    g["b1"] = d1.sum(0)
    flat = np.concatenate([g[name].reshape(-1) for name, _ in SHAPES])
    return loss, flat
