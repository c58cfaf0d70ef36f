"""FP64 reference of the opening-criterion MLP (8 -> 32 -> 32 -> 1, ReLU).

TEST INFRASTRUCTURE ONLY: the checker for the GPU forward / backward
kernels (tree_amd/csrc/mlp.hip).  The reference code has no learned part;
this network is the project's own, so the oracle is pinned by mathematics:
central finite differences of the loss (tests/test_mlp_oracle.py).

Parameter vector layout (shared with the GPU kernels, float32 there):
    W1[8][32]  row-major (input-major), b1[32],
    W2[32][32] row-major,               b2[32],
    w3[32],                             b3
for 1377 values in total.  Loss: mean squared error of the scalar output.
"""
import numpy as np

NIN, NH = 8, 32
NPAR = NIN * NH + NH + NH * NH + NH + NH + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + NIN * NH].reshape(NIN, NH); o += NIN * NH
    b1 = p[o:o + NH]; o += NH
    W2 = p[o:o + NH * NH].reshape(NH, NH); o += NH * NH
    b2 = p[o:o + NH]; o += NH
    w3 = p[o:o + NH]; o += NH
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def forward(p, X):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    z1 = X @ W1 + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2 + b2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ w3 + b3
    return y, (X, z1, h1, z2, h2)


def loss_grad(p, X, t):
    """Mean squared error and its gradient with respect to the parameter vector."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    y, (X, z1, h1, z2, h2) = forward(p, X)
    t = np.asarray(t, dtype=np.float64)
    B = len(t)
    r = y - t
    loss = float(np.mean(r * r))
    dy = 2.0 * r / B
    gw3 = h2.T @ dy
    gb3 = dy.sum()
    dz2 = np.outer(dy, w3) * (z2 > 0)
    gW2 = h1.T @ dz2
    gb2 = dz2.sum(axis=0)
    dz1 = (dz2 @ W2.T) * (z1 > 0)
    gW1 = X.T @ dz1
    gb1 = dz1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return loss, g


def bf16(a):
    """Round to the nearest bfloat16, ties to even, returned as float64 (finite inputs)."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(p, X):
    """The walk's inference arithmetic (matrix cores): bf16 parameters, inputs
    and first hidden layer; exact products; sums in higher precision."""
    W1, b1, W2, b2, w3, b3 = unpack(bf16(p))
    h1 = bf16(np.maximum(bf16(X) @ W1 + b1, 0.0))
    h2 = np.maximum(h1 @ W2 + b2, 0.0)
    return h2 @ w3 + b3


def init_params(seed=0):
    """He-uniform weights, zero biases (the GPU trainer starts from the same vector)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPAR)
    o = 0
    for fan_in, fan_out in ((NIN, NH), (NH, NH)):
        lim = np.sqrt(6.0 / fan_in)
        p[o:o + fan_in * fan_out] = rng.uniform(-lim, lim, fan_in * fan_out)
        o += fan_in * fan_out + fan_out
    lim = np.sqrt(6.0 / NH)
    p[o:o + NH] = rng.uniform(-lim, lim, NH)
    return p
