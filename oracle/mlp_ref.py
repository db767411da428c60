"""FP64 numpy restatement of the selector MLP -- TEST INFRASTRUCTURE ONLY.

16 -> 64 -> 64 -> 2, exact GELU x*Phi(x), mean-squared-error loss over all
B x 2 outputs; parameter layout W1 | b1 | W2 | b2 | W3 | b3 with W row-major
[in][out], as the device kernels use.  Pinned by central finite differences
(tests/test_mlp_oracle.py).  The MLP has no counterpart in the reference.
"""
import numpy as np
from scipy.special import erf

D_IN, H, D_OUT = 16, 64, 2
SIZES = [D_IN * H, H, H * H, H, H * D_OUT, D_OUT]
NPARAMS = sum(SIZES)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = [], 0
    for s in SIZES:
        out.append(p[o:o + s])
        o += s
    W1, b1, W2, b2, W3, b3 = out
    return W1.reshape(D_IN, H), b1, W2.reshape(H, H), b2, W3.reshape(H, D_OUT), b3


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    z1 = X @ W1 + b1
    a1 = gelu(z1)
    z2 = a1 @ W2 + b2
    a2 = gelu(z2)
    return a2 @ W3 + b3, (X, z1, a1, z2, a2)


def loss(p, X, T):
    Y, _ = forward(p, np.asarray(X, np.float64))
    return float(np.mean((Y - T) ** 2))


def loss_grad(p, X, T):
    X = np.asarray(X, np.float64)
    T = np.asarray(T, np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(p)
    Y, (X, z1, a1, z2, a2) = forward(p, X)
    n = Y.size
    dY = 2.0 * (Y - T) / n
    gW3 = a2.T @ dY
    gb3 = dY.sum(0)
    d2 = (dY @ W3.T) * gelu_grad(z2)
    gW2 = a1.T @ d2
    gb2 = d2.sum(0)
    d1 = (d2 @ W2.T) * gelu_grad(z1)
    gW1 = X.T @ d1
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return float(np.mean((Y - T) ** 2)), g
