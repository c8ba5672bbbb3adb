"""The learned space adapter's network (no reference counterpart: NumCME.jl's adapters are
rule-based).

``AdapterMlp`` owns an 8 -> 64 -> 64 -> 1 GELU network on the device whose forward and
backward passes are the library's BF16 MFMA kernels (FP32 accumulation, FP32 master
weights, Adam). Inputs are ``(n, 8)`` feature rows; the output is one keep/add logit per row.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import ArgumentError

N_FEATURES = 8
N_PARAMS = 4801


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _rows(X):
    X = _f32(X)
    if X.ndim != 2 or X.shape[1] != N_FEATURES:
        raise ArgumentError(L.NC_ERR_ARG, f"features must be an (n, {N_FEATURES}) array")
    return X


class AdapterMlp(L.Handle):
    _destroy = "nc_mlp_destroy"

    def __init__(self, seed=0):
        h = C.c_void_p()
        L.call("nc_mlp_create", C.c_uint64(int(seed)), C.byref(h))
        super().__init__(h.value)

    @property
    def params(self):
        out = np.empty(N_PARAMS, dtype=np.float32)
        L.call("nc_mlp_get_params", self.handle, L.ptr(out, C.c_float))
        return out

    @params.setter
    def params(self, p):
        p = _f32(p)
        if p.shape != (N_PARAMS,):
            raise ArgumentError(L.NC_ERR_ARG, f"expected {N_PARAMS} parameters")
        L.call("nc_mlp_set_params", self.handle, L.ptr(p, C.c_float))

    def forward(self, X):
        X = _rows(X)
        out = np.empty(X.shape[0], dtype=np.float32)
        L.call("nc_mlp_forward", self.handle, X.shape[0], L.ptr(X, C.c_float), L.ptr(out, C.c_float))
        return out

    def grad(self, X, Y):
        """Gradient of the mean binary cross-entropy over the rows, and the loss."""
        X, Y = _rows(X), _f32(Y)
        if Y.shape != (X.shape[0],):
            raise ArgumentError(L.NC_ERR_ARG, "one label per row")
        g = np.empty(N_PARAMS, dtype=np.float32)
        loss = C.c_double()
        L.call("nc_mlp_grad", self.handle, X.shape[0], L.ptr(X, C.c_float), L.ptr(Y, C.c_float), L.ptr(g, C.c_float),
               C.byref(loss))
        return g, loss.value

    def train(self, X, Y, epochs=10, batch=4096, lr=1e-3, seed=0):
        """Adam over shuffled mini-batches; returns the mean loss of the last epoch."""
        X, Y = _rows(X), _f32(Y)
        if Y.shape != (X.shape[0],):
            raise ArgumentError(L.NC_ERR_ARG, "one label per row")
        loss = C.c_double()
        L.call("nc_mlp_train", self.handle, X.shape[0], L.ptr(X, C.c_float), L.ptr(Y, C.c_float), int(epochs),
               int(batch), C.c_float(lr), C.c_uint64(int(seed)), C.byref(loss))
        return loss.value

    def bench(self, n=1 << 21, reps=20, train=False):
        """Average kernel duration (ms) of one forward (or forward+backward) launch over n
        synthetic device-resident rows."""
        ms = C.c_double()
        L.call("nc_mlp_bench", self.handle, int(n), int(reps), 1 if train else 0, C.byref(ms))
        return ms.value
