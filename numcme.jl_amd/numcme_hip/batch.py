"""Batched FSP solves (no reference counterpart: NumCME.jl solves one model per call).

``BatchFsp(matrices)`` copies many time-invariant generators (``FspMatrixSparse``) of
uneven size into one device layout; ``solve(u0s, ts)`` integrates all of them together
with Dormand-Prince 5(4), each instance with its own step control, and returns one list of
snapshots (``len(ts)`` arrays of ``rowcount`` entries) per instance.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import ArgumentError


class BatchFsp(L.Handle):
    _destroy = "nc_batch_destroy"

    def __init__(self, matrices):
        self.matrices = list(matrices)
        arr = (C.c_void_p * len(self.matrices))(*[m.handle for m in self.matrices])
        h = C.c_void_p()
        L.call("nc_batch_create", len(self.matrices), arr, C.byref(h))
        self._init_handle(h.value, len(self.matrices))

    @classmethod
    def from_models(cls, models, states, precision="fp32", workers=0):
        """The batch of the generators of ``models[i]`` on the explicit state lists
        ``states[i]`` (rows of species counts), built by worker threads of the library
        (``nc_batch_create_models``) without a matrix object per instance."""
        models, states = list(models), [np.ascontiguousarray(np.asarray(s, dtype=np.int32)) for s in states]
        if len(models) != len(states) or not models:
            raise ArgumentError(L.NC_ERR_ARG, "one state list per model, at least one instance")
        for m, st in zip(models, states):
            if st.ndim != 2 or st.shape[1] != m.get_species_count():
                raise ArgumentError(L.NC_ERR_ARG, "each state list must have one column per species")
        n = len(models)
        hm = (C.c_void_p * n)(*[m.device_model().handle for m in models])
        hs = (C.c_void_p * n)(*[st.ctypes.data for st in states])
        counts = np.array([len(st) for st in states], dtype=np.int64)
        prec = L.NC_FP32 if precision == "fp32" else L.NC_FP64
        h = C.c_void_p()
        self = cls.__new__(cls)
        self.matrices = []
        L.call("nc_batch_create_models", n, hm, hs, L.ptr(counts, C.c_int64), prec, int(workers), C.byref(h))
        self._init_handle(h.value, n)
        return self

    def _init_handle(self, handle, n):
        L.Handle.__init__(self, handle)
        self.n_instances = n
        total = C.c_int64()
        rows = np.empty(n, dtype=np.int64)
        L.call("nc_batch_info", self.handle, C.byref(total), L.ptr(rows, C.c_int64))
        self.total_rows = total.value
        self.rowcounts = rows
        self.stats = None

    def __len__(self):
        return self.n_instances

    def autotune(self, reps=3):
        """Pick the kernels' rows per thread (1, 2 or 4) by timing batched iterations;
        returns (rows_per_thread, ms per iteration)."""
        r, ms = C.c_int32(), C.c_double()
        L.call("nc_batch_autotune", self.handle, int(reps), C.byref(r), C.byref(ms))
        return r.value, ms.value

    def set_rows_per_thread(self, rows):
        L.call("nc_batch_set_rows_per_thread", self.handle, int(rows))

    def solve(self, u0s, ts, rtol=1e-8, atol=1e-14, h0=0.0, graphs=True, fetch=True):
        """Integrate instance i from ``u0s[i]`` through the shared output times ``ts``."""
        if len(u0s) != len(self):
            raise ArgumentError(L.NC_ERR_ARG, "one initial vector per instance")
        for u, n in zip(u0s, self.rowcounts):
            if len(u) != n:
                raise ArgumentError(L.NC_ERR_ARG, "initial vector length must equal the instance's rowcount")
        u0 = L.f64(np.concatenate([np.asarray(u, dtype=np.float64) for u in u0s]))
        ts = L.f64(ts)
        o = L.nc_solve_opts()
        L.call("nc_solve_opts_default", C.byref(o))
        o.method = L.NC_METHOD_RK45
        o.rtol, o.atol, o.h0 = float(rtol), float(atol), float(h0)
        o.no_graphs = 0 if graphs else 1
        st = L.nc_solve_stats()
        L.call("nc_batch_solve", self.handle, C.byref(o), L.ptr(u0, C.c_double), len(ts), L.ptr(ts, C.c_double),
               C.byref(st))
        self.stats = {f: getattr(st, f) for f, _ in L.nc_solve_stats._fields_}
        if not fetch:
            return None
        per_k = [self.snapshots(k) for k in range(len(ts))]
        return [[per_k[k][i] for k in range(len(ts))] for i in range(len(self))]

    def snapshots(self, k):
        """Snapshot k of every instance, downloaded at once: one array per instance (views
        of one buffer)."""
        counts = np.asarray(self.rowcounts, dtype=np.int64)
        out = np.empty(int(counts.sum()))
        L.call("nc_batch_get_snapshots", self.handle, int(k), L.ptr(out, C.c_double))
        return np.split(out, np.cumsum(counts)[:-1])

    def snapshot(self, inst, k):
        out = np.empty(int(self.rowcounts[inst]))
        L.call("nc_batch_get_snapshot", self.handle, int(inst), int(k), L.ptr(out, C.c_double))
        return out
