"""CPU restatement of the autotuner's learned cost model, in float64.

TEST INFRASTRUCTURE ONLY: tests/ use it to check the HIP forward/backward
kernels of pqc_amd/csrc/costmodel.hip; the library never calls it.

Network: in -> 32 -> 32 -> 1, ReLU, loss = mean((y - t)^2).  Parameters are
flattened as W1[H][D], b1[H], W2[H][H], b2[H], w3[H], b3, the layout the HIP
kernels use.  The gradient is written out by hand (backpropagation) and pinned
against central finite differences in tests/test_costmodel_cpu.py.
"""
import numpy as np

H = 32


def n_params(d):
    return H * d + H + H * H + H + H + 1


def unpack(p, d):
    p = np.asarray(p, dtype=np.float64)
    i = 0
    W1 = p[i:i + H * d].reshape(H, d); i += H * d
    b1 = p[i:i + H]; i += H
    W2 = p[i:i + H * H].reshape(H, H); i += H * H
    b2 = p[i:i + H]; i += H
    w3 = p[i:i + H]; i += H
    b3 = p[i]
    return W1, b1, W2, b2, w3, b3


def forward(p, x):
    x = np.asarray(x, dtype=np.float64)
    W1, b1, W2, b2, w3, b3 = unpack(p, x.shape[1])
    z1 = x @ W1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    return h2 @ w3 + b3, (z1, h1, z2, h2)


def loss(p, x, t):
    y, _ = forward(p, x)
    return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))


def grad(p, x, t):
    """(loss, dloss/dp) by backpropagation."""
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    n, d = x.shape
    W1, b1, W2, b2, w3, b3 = unpack(p, d)
    y, (z1, h1, z2, h2) = forward(p, x)
    e = y - t
    dy = 2.0 * e / n
    gw3 = h2.T @ dy
    gb3 = dy.sum()
    d2 = np.outer(dy, w3) * (z2 > 0)
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (z1 > 0)
    gW1 = d1.T @ x
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return float(np.mean(e ** 2)), g


def adam_fit(p, x, t, steps, lr, b1=0.9, b2=0.999, eps=1e-8):
    """Full-batch Adam, the update rule of k_mlp_adam."""
    p = np.array(p, dtype=np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for k in range(1, steps + 1):
        _, g = grad(p, x, t)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        p -= lr * (m / (1 - b1 ** k)) / (np.sqrt(v / (1 - b2 ** k)) + eps)
    return p
