"""Learned space adapter support (no reference counterpart: NumCME.jl's adapters are
rule-based).

* ``boundary_features(space, model, p, dsinks, t, tend)`` — the device feature rows of the
  boundary states the selective adapter would expand from (ids 0-based, ascending).
* ``TrainingRecords`` — (features, label) rows that an adaptive solve collects when passed
  as ``records=``: each adaptation records up to ``per_adaptation`` boundary states and
  labels each one at the next adaptation (or at the end of the solve) with 1 when its
  successors along its sink reactions then hold at least ``label_mass`` probability.
* ``train_adapter(X, Y, ...)`` — an ``AdapterMlp`` trained on such rows on the device.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import ArgumentError
from .mlp import N_FEATURES, AdapterMlp


def boundary_features(space, model, p, dsinks, t, tend):
    p = L.f64(p)
    ds = L.f64(dsinks)
    n = space.get_state_count()
    if p.size != n:
        raise ArgumentError(L.NC_ERR_ARG, "p must hold one probability per state")
    if ds.size != space.get_sink_count():
        raise ArgumentError(L.NC_ERR_ARG, "dsinks must have one entry per reaction")
    ids = np.empty(max(n, 1), dtype=np.int64)
    F = np.empty((max(n, 1), N_FEATURES), dtype=np.float32)
    nb = C.c_int64()
    L.call("nc_adapter_features", space.handle, model.device_model().handle, n,
           L.ptr(p if n else np.zeros(1), C.c_double), L.ptr(ds, C.c_double), float(t), float(tend), C.byref(nb),
           L.ptr(ids, C.c_int64), L.ptr(F, C.c_float))
    return ids[: nb.value].copy(), F[: nb.value].copy()


class TrainingRecords(L.Handle):
    _destroy = "nc_records_destroy"

    def __init__(self, capacity=1_000_000, per_adaptation=4096, label_mass=1e-9):
        h = C.c_void_p()
        L.call("nc_records_create", int(capacity), int(per_adaptation), float(label_mass), C.byref(h))
        super().__init__(h.value)
        self.capacity = int(capacity)

    def __len__(self):
        n = C.c_int64()
        L.call("nc_records_count", self.handle, C.byref(n))
        return n.value

    def arrays(self):
        """(features (n, 8) float32, labels (n,) float32) of the labelled records."""
        n = len(self)
        X = np.empty((max(n, 1), N_FEATURES), dtype=np.float32)
        Y = np.empty(max(n, 1), dtype=np.float32)
        L.call("nc_records_get", self.handle, L.ptr(X, C.c_float), L.ptr(Y, C.c_float))
        return X[:n], Y[:n]

    def clear(self):
        L.call("nc_records_clear", self.handle)


def train_adapter(X, Y, epochs=20, batch=4096, lr=2e-3, seed=0, mlp=None):
    """Train (or continue training) an ``AdapterMlp`` on records; returns (mlp, last loss)."""
    net = AdapterMlp(seed) if mlp is None else mlp
    loss = net.train(X, Y, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return net, loss
