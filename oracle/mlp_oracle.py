"""float64 numpy restatement of the engine's cost-model MLP (4 -> 64 -> 64 ->
K, ReLU, masked mean-squared error) used to check the HIP gradients.

TEST INFRASTRUCTURE ONLY.  The parameter layout matches mlp.hip:
W1[64,4], b1[64], W2[64,64], b2[64], W3[K,64], b3[K], flattened in that
order.  The reference has no learned component; this oracle is pinned by
central finite differences (tests/test_oracle.py).
"""
import numpy as np

H, IN = 64, 4


def unpack(p, K):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + H * IN].reshape(H, IN); o += H * IN
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + K * H].reshape(K, H); o += K * H
    b3 = p[o:o + K]; o += K
    assert o == p.size
    return W1, b1, W2, b2, W3, b3


def forward(p, X, K):
    W1, b1, W2, b2, W3, b3 = unpack(p, K)
    X = np.asarray(X, dtype=np.float64)
    a1 = X @ W1.T + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0)
    return h2 @ W3.T + b3, (X, a1, h1, a2, h2)


def loss(p, X, T, M, K):
    Y, _ = forward(p, X, K)
    M = np.asarray(M, dtype=np.float64)
    return float((M * (Y - T) ** 2).sum() / M.sum())


def gradients(p, X, T, M, K):
    W1, b1, W2, b2, W3, b3 = unpack(p, K)
    Y, (X, a1, h1, a2, h2) = forward(p, X, K)
    M = np.asarray(M, dtype=np.float64)
    dY = M * 2.0 * (Y - T) / M.sum()
    gW3 = dY.T @ h2
    gb3 = dY.sum(0)
    dh2 = (dY @ W3) * (a2 > 0)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2) * (a1 > 0)
    gW1 = dh1.T @ X
    gb1 = dh1.sum(0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3]), loss(p, X, T, M, K)
