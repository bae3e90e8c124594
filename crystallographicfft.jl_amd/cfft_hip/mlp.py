"""Warm-start network on the GPU (C context ``cfft_mlp``): a multilayer
perceptron trained and evaluated by the library's own kernels (BF16 matrix-core
GEMMs with FP32 accumulation, SiLU, MSE, Adam).  No counterpart in the
reference."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def _fp(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(L.PF)


class MLP:
    def __init__(self, dims, seed=0):
        self.dims = [int(d) for d in dims]
        h = C.c_void_p()
        arr = np.array(self.dims, dtype=np.int32)
        L.check(L.lib().cfft_mlp_create(len(self.dims) - 1, L.i32p(arr), int(seed), C.byref(h)))
        self._h = h
        n = C.c_int64()
        L.check(L.lib().cfft_mlp_info(h, C.byref(n)))
        self.n_params = int(n.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                L.lib().cfft_mlp_destroy(h)
            except Exception:
                pass
            self._h = None

    def params(self):
        out = np.zeros(self.n_params, dtype=np.float32)
        L.check(L.lib().cfft_mlp_get_params(self._h, _fp(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.n_params:
            raise L.CFFTError(L.CFFT_ERR_INVALID, "parameter count mismatch")
        L.check(L.lib().cfft_mlp_set_params(self._h, _fp(p)))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.dims[0])
        y = np.zeros((x.shape[0], self.dims[-1]), dtype=np.float32)
        L.check(L.lib().cfft_mlp_forward(self._h, _fp(x), x.shape[0], _fp(y)))
        return y

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.dims[0])
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, self.dims[-1])
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = C.c_double()
        L.check(L.lib().cfft_mlp_gradient(self._h, _fp(x), _fp(t), x.shape[0], C.byref(loss), _fp(g)))
        return float(loss.value), g

    def train(self, X, T, epochs, batch=64, lr=1e-3, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, self.dims[0])
        T = np.ascontiguousarray(T, dtype=np.float32).reshape(-1, self.dims[-1])
        hist = np.zeros(max(epochs, 1))
        L.check(L.lib().cfft_mlp_train(self._h, _fp(X), _fp(T), X.shape[0], int(epochs), int(batch),
                                       float(lr), int(seed), hist.ctypes.data_as(L.PD)))
        return hist[:epochs]
