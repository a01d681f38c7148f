"""FP64 numpy restatement of the surrogate network, for gradient checks.

TEST INFRASTRUCTURE ONLY.  Same parameter layout as the HIP kernels:
W1[64][16] b1[64] W2[64][64] b2[64] W3[2][64] b3[2]; exact-erf GELU;
loss per sample 0.5*(s + (y - mu)^2 exp(-s)), mean over the batch.
The surrogate has no counterpart in the reference; these checks are pinned
by finite differences (tests/test_surrogate_cpu.py).
"""
import numpy as np
from scipy.special import erf

NPARAMS = 64 * 16 + 64 + 64 * 64 + 64 + 2 * 64 + 2


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + 1024].reshape(64, 16); o += 1024
    b1 = p[o:o + 64]; o += 64
    W2 = p[o:o + 4096].reshape(64, 64); o += 4096
    b2 = p[o:o + 64]; o += 64
    W3 = p[o:o + 128].reshape(2, 64); o += 128
    b3 = p[o:o + 2]
    return W1, b1, W2, b2, W3, b3


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2 * np.pi)


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    z1 = X @ W1.T + b1
    h1 = gelu(z1)
    z2 = h1 @ W2.T + b2
    h2 = gelu(z2)
    o = h2 @ W3.T + b3
    return o, (z1, h1, z2, h2)


def loss(p, X, y):
    o, _ = forward(p, X)
    mu, s = o[:, 0], o[:, 1]
    return float(np.mean(0.5 * (s + (y - mu) ** 2 * np.exp(-s))))


def grad(p, X, y):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    o, (z1, h1, z2, h2) = forward(p, X)
    B = X.shape[0]
    mu, s = o[:, 0], o[:, 1]
    r = y - mu
    e = np.exp(-s)
    d3 = np.stack([-r * e, 0.5 * (1 - r * r * e)], axis=1) / B
    dW3 = d3.T @ h2
    db3 = d3.sum(0)
    d2 = (d3 @ W3) * gelu_grad(z2)
    dW2 = d2.T @ h1
    db2 = d2.sum(0)
    d1 = (d2 @ W2) * gelu_grad(z1)
    dW1 = d1.T @ X
    db1 = d1.sum(0)
    return np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3.ravel(), db3])


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    parts = [rng.standard_normal((64, 16)) * np.sqrt(1 / 16), np.zeros(64),
             rng.standard_normal((64, 64)) * np.sqrt(1 / 64), np.zeros(64),
             rng.standard_normal((2, 64)) * np.sqrt(1 / 64) * 0.1, np.zeros(2)]
    return np.concatenate([q.ravel() for q in parts])
