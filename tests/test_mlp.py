"""CPU checks that pin the cost-model oracle (oracle/mlp_ref.py)."""
import numpy as np

import mlp_ref


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.140625, 0.0], np.float32)
    r = mlp_ref.bf16_round(x)
    # 1 + 2^-8 is halfway between bf16 neighbours 1 and 1 + 2^-7: ties to even -> 1
    # 1 + 3*2^-9 is above the halfway point -> 1 + 2^-7
    assert r.tolist() == [1.0, 1.0, 1.0 + 2 ** -7, -2.5, 3.140625, 0.0]
    assert mlp_ref.P == 1410


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = mlp_ref.bf16_round(rng.normal(0, 0.4, mlp_ref.P).astype(np.float32))
    X = rng.normal(0, 1, (40, 8))
    T = rng.normal(0, 1, (40, 2))
    M = (rng.random((40, 2)) > 0.3).astype(np.float64)
    loss, g = mlp_ref.loss_grad(p, X, T, M, dtype=np.float64)
    # perturb the (already bf16-exact) parameters in float64, no re-rounding
    def L(q):
        W = {}
        o = 0
        for name, shape in mlp_ref.SHAPES:
            k = int(np.prod(shape))
            W[name] = q[o:o + k].reshape(shape)
            o += k
        h1 = np.maximum(X @ W["W1"] + W["b1"], 0)
        h2 = np.maximum(h1 @ W["W2"] + W["b2"], 0)
        y = h2 @ W["W3"] + W["b3"]
        return (M * (y - T) ** 2).sum() / M.sum()
    q = p.astype(np.float64)
    assert abs(L(q) - loss) < 1e-12
    eps = 1e-6
    for i in rng.choice(mlp_ref.P, 60, replace=False).tolist() + [mlp_ref.P - 1, mlp_ref.P - 2]:
        qp, qm = q.copy(), q.copy()
        qp[i] += eps
        qm[i] -= eps
        fd = (L(qp) - L(qm)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_oracle_planted_mistake_is_caught():
    # a dropped ReLU mask in the backward pass must fail the check above
    rng = np.random.default_rng(1)
    p = mlp_ref.bf16_round(rng.normal(0, 0.4, mlp_ref.P).astype(np.float32))
    X = rng.normal(0, 1, (30, 8))
    T = rng.normal(0, 1, (30, 2))
    M = np.ones((30, 2))
    W = mlp_ref.unpack(p, np.float64)
    y, (x, h1, h2) = mlp_ref.forward(p, X, dtype=np.float64)
    dy = 2 * (y - T) / M.sum()
    wrong = h1.T @ (dy @ W["W3"].T)          # no (h2 > 0) mask
    _, g = mlp_ref.loss_grad(p, X, T, M, dtype=np.float64)
    right = g[mlp_ref.P - 66 - 32 - 1024: mlp_ref.P - 66 - 32].reshape(32, 32)
    assert not np.allclose(wrong, right)
