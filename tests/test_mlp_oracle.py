"""The FP64 MLP oracle (oracle/mlp_ref.py) pinned by central finite
differences of its own loss, and by a hand-computed tiny case."""
import numpy as np

from oracle import mlp_ref


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = mlp_ref.init_params(3)
    X = rng.normal(size=(37, 8))
    t = rng.normal(size=37)
    loss, g = mlp_ref.loss_grad(p, X, t)
    h = 1e-6
    idx = rng.choice(mlp_ref.NPAR, 120, replace=False)
    idx = np.r_[idx, [mlp_ref.NPAR - 1, 0, 8 * 32, 8 * 32 + 32]]  # b3, W1[0,0], b1[0], W2[0,0]
    for j in idx:
        pp, pm = p.copy(), p.copy()
        pp[j] += h
        pm[j] -= h
        fd = (mlp_ref.loss_grad(pp, X, t)[0] - mlp_ref.loss_grad(pm, X, t)[0]) / (2 * h)
        assert abs(fd - g[j]) <= 1e-6 * max(1.0, abs(fd)), (j, fd, g[j])


def test_tiny_network_by_hand():
    """One active path: y = w3 * relu(W2 * relu(W1 x + b1) + b2) + b3."""
    p = np.zeros(mlp_ref.NPAR)
    W1, b1, W2, b2, w3, b3 = [a for a in mlp_ref.unpack(p)]
    o = 0
    p[0 * 32 + 0] = 2.0           # W1[0,0]
    o = 8 * 32
    p[o + 0] = -1.0               # b1[0]
    o += 32
    p[o + 0 * 32 + 0] = 3.0       # W2[0,0]
    o += 32 * 32
    p[o + 0] = 0.5                # b2[0]
    o += 32
    p[o + 0] = 4.0                # w3[0]
    p[-1] = 0.25                  # b3
    x = np.zeros((1, 8))
    x[0, 0] = 1.5
    y, _ = mlp_ref.forward(p, x)
    h1 = max(2.0 * 1.5 - 1.0, 0)
    h2 = max(3.0 * h1 + 0.5, 0)
    assert y[0] == 4.0 * h2 + 0.25
    x[0, 0] = 0.25  # h1 < 0 -> relu cuts the path
    y, _ = mlp_ref.forward(p, x)
    assert y[0] == 4.0 * 0.5 + 0.25
