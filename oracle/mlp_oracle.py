"""numpy restatement of the stage-guess network's forward, backward and Adam
(TEST INFRASTRUCTURE ONLY; imported by tests/ as the checker of mlp.hip).

No reference counterpart: the reference's stage guesses are zero
(openmp_impl/schrodinger.hpp:222 etc.). The GPU rounds both GEMM operands to
bfloat16 (round to nearest even) and accumulates in FP32; this restatement
rounds the same operands the same way and accumulates in float64, storing
layer outputs as float32 like the GPU does. GELU is the erf form
0.5 x (1 + erf(x / sqrt 2)).
"""
import numpy as np
from scipy.special import erf


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p, dims, in_pad, out_pad, off):
    n_in, h1, h2, n_out = dims
    W1 = p[off[0]:off[1]].reshape(in_pad, h1)
    b1 = p[off[1]:off[2]]
    W2 = p[off[2]:off[3]].reshape(h1, h2)
    b2 = p[off[3]:off[4]]
    W3 = p[off[4]:off[5]].reshape(h2, out_pad)
    b3 = p[off[5]:off[5] + out_pad]
    return W1, b1, W2, b2, W3, b3


def gelu(x):
    return 0.5 * x * (1 + erf(x / np.sqrt(2)))


def gelu_d(x):
    return 0.5 * (1 + erf(x / np.sqrt(2))) + x * np.exp(-0.5 * x * x) / np.sqrt(2 * np.pi)


def mm(a, b):
    return (bf16(a) @ bf16(b)).astype(np.float32)


def forward(p, X, dims, in_pad, out_pad, off):
    W1, b1, W2, b2, W3, b3 = unpack(p, dims, in_pad, out_pad, off)
    H1 = (mm(X, W1) + b1).astype(np.float32)
    A1 = gelu(H1.astype(np.float64)).astype(np.float32)
    H2 = (mm(A1, W2) + b2).astype(np.float32)
    A2 = gelu(H2.astype(np.float64)).astype(np.float32)
    Y = (mm(A2, W3) + b3).astype(np.float32)
    return Y, (H1, A1, H2, A2)


def gradients(p, X, T, dims, in_pad, out_pad, off):
    """MSE loss mean over rows x real outputs; returns (loss, flat gradient)."""
    n_in, h1, h2, n_out = dims
    W1, b1, W2, b2, W3, b3 = unpack(p, dims, in_pad, out_pad, off)
    Y, (H1, A1, H2, A2) = forward(p, X, dims, in_pad, out_pad, off)
    rows = X.shape[0]
    D = np.zeros_like(Y, dtype=np.float64)
    D[:, :n_out] = Y[:, :n_out].astype(np.float64) - T[:, :n_out]
    loss = float(np.sum(D[:, :n_out] ** 2) / (rows * n_out))
    dY = (2.0 / (rows * n_out) * D).astype(np.float32)
    g = np.zeros_like(p, dtype=np.float64)
    gW3 = bf16(A2).T @ bf16(dY)
    gb3 = dY.astype(np.float64).sum(0)
    dA2 = (mm(dY, W3.T) * gelu_d(H2.astype(np.float64))).astype(np.float32)
    gW2 = bf16(A1).T @ bf16(dA2)
    gb2 = dA2.astype(np.float64).sum(0)
    dA1 = (mm(dA2, W2.T) * gelu_d(H1.astype(np.float64))).astype(np.float32)
    gW1 = bf16(X).T @ bf16(dA1)
    gb1 = dA1.astype(np.float64).sum(0)
    for o, v in zip(off, (gW1, gb1, gW2, gb2, gW3, gb3)):
        g[o:o + v.size] = v.ravel()
    return loss, g


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """Kingma & Ba (2015) with bias correction, in float32 like the GPU."""
    p, g, m, v = (np.asarray(a, dtype=np.float32) for a in (p, g, m, v))
    m = np.float32(b1) * m + np.float32(1 - b1) * g
    v = np.float32(b2) * v + np.float32(1 - b2) * g * g
    c1 = np.float32(1 - b1 ** t)
    c2 = np.float32(1 - b2 ** t)
    p = p - np.float32(lr) * (m / c1) / (np.sqrt(v / c2) + np.float32(eps))
    return p, m, v
