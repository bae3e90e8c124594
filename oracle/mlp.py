"""TEST INFRASTRUCTURE ONLY: NumPy restatement of the warm-start network.

Same parameter layout as the library (W_0 .. W_{L-1} row-major dims[l] x
dims[l+1], then the biases), SiLU hidden layers, linear output, mean-squared
error.  GEMM operands are rounded to BF16 (round to nearest even) exactly as
the device loads them; products are accumulated in FP64 here (FP32 on the
device), so gradients agree to FP32 accumulation error.
"""
from __future__ import annotations

import numpy as np

__all__ = ["bf16", "unpack", "forward", "loss_and_grad"]


def bf16(x):
    """Round FP32 values to BF16 (nearest even), returned as FP32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(params, dims):
    L = len(dims) - 1
    W, B = [], []
    off = 0
    for l in range(L):
        n = dims[l] * dims[l + 1]
        W.append(params[off:off + n].reshape(dims[l], dims[l + 1]))
        off += n
    for l in range(L):
        B.append(params[off:off + dims[l + 1]])
        off += dims[l + 1]
    return W, B


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def _mm(a, b, rounding):
    if not rounding:
        return np.asarray(a, dtype=np.float64) @ np.asarray(b, dtype=np.float64)
    return bf16(a).astype(np.float64) @ bf16(b).astype(np.float64)


def forward(params, dims, x, rounding=True):
    """rounding=False: plain FP64 (used to pin the gradient by finite
    differences); rounding=True: the device's BF16 operands."""
    ft = np.float32 if rounding else np.float64
    W, B = unpack(np.asarray(params, dtype=ft), dims)
    h = np.asarray(x, dtype=ft)
    zs, hs = [], [h]
    for l in range(len(W)):
        z = (_mm(h, W[l], rounding) + B[l]).astype(ft)
        zs.append(z)
        h = _silu(z).astype(ft) if l + 1 < len(W) else z
        hs.append(h)
    return zs, hs


def loss_and_grad(params, dims, x, t, rounding=True):
    ft = np.float32 if rounding else np.float64
    W, B = unpack(np.asarray(params, dtype=ft), dims)
    zs, hs = forward(params, dims, x, rounding)
    y = zs[-1]
    n = y.size
    loss = float(np.mean((y.astype(np.float64) - t) ** 2))
    dz = (2.0 * (y - t) / n).astype(ft)
    gW, gB = [None] * len(W), [None] * len(W)
    for l in reversed(range(len(W))):
        gW[l] = _mm(hs[l].T, dz, rounding).astype(ft)
        gB[l] = dz.sum(axis=0).astype(ft)
        if l == 0:
            break
        dz = (_mm(dz, W[l].T, rounding) * _dsilu(zs[l - 1])).astype(ft)
    return loss, np.concatenate([g.ravel() for g in gW] + [g.ravel() for g in gB])
