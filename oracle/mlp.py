"""CPU oracle of the learned warm start's network — TEST INFRASTRUCTURE ONLY.

A numpy restatement of the MLP the library trains with its own kernels
(megastation_amd/csrc/mlp.hip, sequenced by megastation_amd/learned.py):
layers z = h W + b, GELU (tanh form) on the hidden layers, identity on the
output; MSE loss 0.5/M * sum |y - t|^2; Adam. With bf16=True the operands are
rounded to BF16 where the kernels store BF16 (weights, hidden activations,
the backward's dZ) and everything else is computed in FP64, so the kernels'
results differ from it only by FP32 accumulation order and the rare BF16
rounding flips that follow. The network has no counterpart in the reference;
this oracle is pinned by central finite differences of its own loss
(tests/test_oracle_mlp.py).
"""
from __future__ import annotations

import numpy as np

K0, K1 = 0.7978845608028654, 0.044715


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    out = (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(np.isfinite(f), out, f.astype(np.float64))


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(K0 * (z + K1 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(K0 * (z + K1 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * K0 * (1.0 + 3.0 * K1 * z * z)


def forward(x, Ws, bs, rnd: bool = True):
    """x [M][d0]; returns (out FP64 [M][d3], cache)."""
    R = bf16 if rnd else (lambda a: np.asarray(a, dtype=np.float64))
    h = R(x)
    hs, zs = [h], []
    for l, (W, b) in enumerate(zip(Ws, bs)):
        z = h @ R(W) + np.asarray(b, dtype=np.float64)
        zs.append(z)
        if l < len(Ws) - 1:
            h = R(gelu(z))
            hs.append(h)
    return zs[-1], (hs, zs)


def backward(dout, Ws, cache, rnd: bool = True):
    """Gradients (dWs, dbs) for upstream dout [M][d3]."""
    R = bf16 if rnd else (lambda a: np.asarray(a, dtype=np.float64))
    hs, zs = cache
    L = len(Ws)
    dWs, dbs = [None] * L, [None] * L
    dy = np.asarray(dout, dtype=np.float64)
    for l in reversed(range(L)):
        dz = dy if l == L - 1 else dy * gelu_grad(zs[l])
        dzb = R(dz)
        dWs[l] = hs[l].T @ dzb
        dbs[l] = dz.sum(axis=0)
        if l > 0:
            dy = dzb @ R(Ws[l]).T
    return dWs, dbs


def mse(out, target):
    M = out.shape[0]
    d = out - target
    return 0.5 * (d * d).sum() / M, d / M


def adam(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v
