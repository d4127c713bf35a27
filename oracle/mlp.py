"""FP64 oracle of the cost model's MLP (numpy).

TEST INFRASTRUCTURE ONLY: checks the device forward/backward kernels of
convgemm_amd/csrc/kernels/mlp.hip.  Network: x (F) -> relu(x W1 + b1) (H)
-> relu(h1 W2 + b2) (H) -> h2 W3 + b3 (1); loss = mean((y - t)^2).
Parameters are one flat vector [W1 (FxH), b1, W2 (HxH), b2, W3 (Hx1), b3],
row-major.  Pinned in tests/test_oracle.py by central finite differences and
against torch's float64 linear layers and autograd.
"""
import numpy as np

F, H = 33, 64
NP = F * H + H + H * H + H + H + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + F * H].reshape(F, H); o += F * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H].reshape(H, 1); o += H
    b3 = p[o:o + 1]
    return W1, b1, W2, b2, W3, b3


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    h1 = np.maximum(X @ W1 + b1, 0.0)
    h2 = np.maximum(h1 @ W2 + b2, 0.0)
    return (h2 @ W3 + b3)[:, 0], (h1, h2)


def loss(p, X, T):
    y, _ = forward(p, X)
    return float(np.mean((y - np.asarray(T, np.float64)) ** 2))


def loss_grad(p, X, T):
    """(loss, flat gradient) by reverse-mode differentiation in float64."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    T = np.asarray(T, np.float64)
    N = X.shape[0]
    y, (h1, h2) = forward(p, X)
    dy = 2.0 * (y - T) / N
    dW3 = h2.T @ dy[:, None]
    db3 = np.array([dy.sum()])
    dz2 = (dy[:, None] @ W3.T) * (h2 > 0)
    dW2 = h1.T @ dz2
    db2 = dz2.sum(0)
    dz1 = (dz2 @ W2.T) * (h1 > 0)
    dW1 = X.T @ dz1
    db1 = dz1.sum(0)
    g = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3.ravel(), db3])
    return float(np.mean((y - T) ** 2)), g
