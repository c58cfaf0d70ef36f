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


def test_bf16_rounding():
    """Round to nearest, ties to even, against exactly representable cases."""
    one = 1.0
    ulp = 2.0 ** -7  # bf16 spacing in [1, 2)
    cases = [(one, one), (one + ulp / 4, one), (one + 3 * ulp / 4, one + ulp),
             (one + ulp / 2, one),                  # tie -> even (1.0)
             (one + 3 * ulp / 2, one + 2 * ulp),    # tie -> even (1 + 2 ulp)
             (-(one + 3 * ulp / 4), -(one + ulp)), (0.0, 0.0), (257.0, 256.0), (259.0, 260.0)]
    for x, want in cases:
        got = float(mlp_ref.bf16(np.float32(x)))
        assert got == np.float32(want), (x, got, want)


def test_forward_bf16_rounds_only_inputs_params_and_first_layer():
    """Small integers are exact in bf16, so parameters, inputs and the first
    hidden layer pass through unchanged; the second layer's odd sums above 256
    are not representable in bf16 and must not be rounded: the bf16 forward then
    equals the FP64 forward exactly.  Non-integer inputs do get rounded."""
    rng = np.random.default_rng(3)
    p = rng.integers(-3, 4, size=mlp_ref.NPAR).astype(np.float64)
    X = rng.integers(-4, 5, size=(64, 8)).astype(np.float64)
    W1, b1, W2, b2, w3, b3 = mlp_ref.unpack(p)
    h1 = np.maximum(X @ W1 + b1, 0.0)
    assert h1.max() < 256  # first layer exact in bf16
    h2 = np.maximum(h1 @ W2 + b2, 0.0)
    assert np.any((h2 > 256) & (h2 % 2 == 1))  # second layer not representable
    assert np.array_equal(mlp_ref.forward_bf16(p, X), mlp_ref.forward(p, X)[0])
    Xf = X + 1.0 / 1024.0  # below bf16 resolution for |x| >= 1: rounded away
    big = np.all(np.abs(X) >= 1, axis=1)
    assert np.array_equal(mlp_ref.forward_bf16(p, Xf)[big], mlp_ref.forward(p, X)[0][big])
