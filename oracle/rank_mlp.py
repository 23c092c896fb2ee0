"""The rank model restated in float64 numpy (test infrastructure only).

No counterpart in the reference: the model is this library's addition
(an MLP that predicts the numerical rank of a node-pair block).  Layout of
the flat parameter vector, as the device kernels hold it: W1 (64 x 6), b1,
W2 (64 x 64), b2, W3 (64 x 64), b3, W4 (1 x 64), b4, each W row-major
W[out][in].  Activation GELU(z) = z (1 + erf(z / sqrt 2)) / 2.  Loss: mean
squared error over the samples.  Features and targets here are already
standardized (the device model standardizes with its stored mean and
deviation before the kernels run).

Pinned in tests/test_oracle.py by central differences of `loss` (the
gradient) and by the closed form of a one-unit case.
"""
import numpy as np
from scipy.special import erf

SHAPES = [(64, 6), (64,), (64, 64), (64,), (64, 64), (64,), (1, 64), (1,)]
NPARAM = sum(int(np.prod(s)) for s in SHAPES)


def unpack(p):
    out, o = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(np.asarray(p[o:o + n], np.float64).reshape(s))
        o += n
    return out


def pack(parts):
    return np.concatenate([np.asarray(a, np.float64).ravel() for a in parts])


def gelu(z):
    return 0.5 * z * (1 + erf(z / np.sqrt(2)))


def gelu_d(z):
    return 0.5 * (1 + erf(z / np.sqrt(2))) + z * np.exp(-0.5 * z * z) / np.sqrt(2 * np.pi)


def forward(p, X):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    X = np.asarray(X, np.float64)
    z1 = X @ W1.T + b1
    a1 = gelu(z1)
    z2 = a1 @ W2.T + b2
    a2 = gelu(z2)
    z3 = a2 @ W3.T + b3
    a3 = gelu(z3)
    return (a3 @ W4.T + b4)[:, 0], (z1, a1, z2, a2, z3, a3)


def loss(p, X, y):
    out, _ = forward(p, X)
    return float(np.mean((out - np.asarray(y, np.float64)) ** 2))


def loss_grad(p, X, y):
    """(mean squared error, its gradient in the flat layout)"""
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    X = np.asarray(X, np.float64)
    out, (z1, a1, z2, a2, z3, a3) = forward(p, X)
    d = out - np.asarray(y, np.float64)
    B = len(d)
    dy = 2 * d / B
    gW4 = (dy @ a3)[None, :]
    gb4 = np.array([dy.sum()])
    dz3 = dy[:, None] * W4[0][None, :] * gelu_d(z3)
    gW3, gb3 = dz3.T @ a2, dz3.sum(0)
    dz2 = (dz3 @ W3) * gelu_d(z2)
    gW2, gb2 = dz2.T @ a1, dz2.sum(0)
    dz1 = (dz2 @ W2) * gelu_d(z1)
    gW1, gb1 = dz1.T @ X, dz1.sum(0)
    return float(np.mean(d * d)), pack([gW1, gb1, gW2, gb2, gW3, gb3, gW4, gb4])


def adam_first_step(p, g, lr, eps=1e-8):
    """Parameters after Adam's first step from zero moments (beta 0.9 /
    0.999, bias-corrected): p - lr g / (|g| + eps)."""
    return p - lr * g / (np.abs(g) + eps)
