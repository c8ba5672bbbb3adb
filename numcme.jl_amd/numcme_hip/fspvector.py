"""Sparse multi-indexed probability vectors — mirrors ``src/fspvector/fspvector.jl``.

``FspVectorSparse`` is the snapshot type ``solve`` returns: a state list (rows of an
int64 array) and values. A snapshot of a solve stays on the GPU (``nc_fspvec``) until its
``states`` or ``values`` are read; the marginal ``sum(p, dims)`` and the dense ``Array(p)``
run on the GPU (``nc_fspvec_marginal``, ``nc_fspvec_dense``) and need no download of the
full vector. A vector built on the host from arrays is uploaded when a device operation
first needs it. The state -> index dictionary is built on the host only when asked for.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import ArgumentError


class _DeviceVector(L.Handle):
    _destroy = "nc_fspvec_destroy"


class FspVectorSparse:
    """``FspVectorSparse(states, values; checksizes=true)`` (fspvector.jl:13-35)."""

    def __init__(self, states, values, checksizes=True, copy=True):
        st = np.array(states, dtype=np.int64, copy=copy)
        if st.ndim == 1:
            st = st.reshape(len(values) if len(values) else 0, -1) if st.size else st.reshape(0, 0)
        vals = np.array(values, dtype=np.float64, copy=copy).ravel()
        if checksizes and st.shape[0] != vals.shape[0]:
            raise ArgumentError(L.NC_ERR_ARG, "State and value lists must have equal lengths.")
        self._states = st
        self._values = vals
        self._ns = st.shape[1] if st.ndim == 2 else 0
        self._n = vals.shape[0]
        self._dev = None
        self._dict = None

    @classmethod
    def _from_device(cls, handle):
        """A vector that lives on the device (``nc_fspvec`` handle, owned from here on)."""
        obj = cls.__new__(cls)
        obj._dev = _DeviceVector(handle)
        n, ns = C.c_int64(), C.c_int32()
        L.call("nc_fspvec_size", handle, C.byref(n), C.byref(ns))
        obj._n, obj._ns = n.value, ns.value
        obj._states = obj._values = None
        obj._dict = None
        return obj

    @classmethod
    def from_space(cls, space, statevalpairs):
        """``FspVectorSparse(statespace, [x => v, ...])`` (fspvector.jl:42-55)."""
        states = space.get_states()
        vals = np.zeros(states.shape[0])
        if statevalpairs:
            keys = np.asarray([k for k, _ in statevalpairs], dtype=np.int64).reshape(len(statevalpairs), -1)
            idx = space.lookup(keys)
            for (_, v), i in zip(statevalpairs, idx):
                if i != 0:
                    vals[i - 1] = v
        return cls(states, vals, copy=False)

    # ---- host view (each part downloaded on its first access) ------------------------------
    @property
    def states(self):
        if self._states is None:
            st = np.empty((self._n, self._ns), dtype=np.int32)
            if st.size:
                L.call("nc_fspvec_get", self._dev.handle, L.ptr(st, C.c_int32), None)
            self._states = st.astype(np.int64)
        return self._states

    @property
    def values(self):
        if self._values is None:
            vals = np.empty(self._n, dtype=np.float64)
            if vals.size:
                L.call("nc_fspvec_get", self._dev.handle, None, L.ptr(vals, C.c_double))
            self._values = vals
        return self._values

    def on_device(self):
        """True while neither the states nor the values have been downloaded to the host."""
        return self._states is None and self._values is None

    def _device(self):
        if self._dev is None:
            st = L.i32(self._states.reshape(self._n, self._ns))
            vals = L.f64(self._values)
            h = C.c_void_p()
            L.call("nc_fspvec_create", self._ns, self._n, L.ptr(st, C.c_int32), L.ptr(vals, C.c_double),
                   C.byref(h))
            self._dev = _DeviceVector(h.value)
        return self._dev.handle

    @property
    def state2idx(self):
        if self._dict is None:
            self._dict = {tuple(x): i + 1 for i, x in enumerate(self.states.tolist())}
        return self._dict

    def get_states(self):
        return self.states

    def get_values(self):
        return self.values

    def get_statedict(self):
        return self.state2idx

    def nnz(self):
        return self._n

    def sum(self, dims=None):
        """``sum(p)`` (fspvector.jl:57-59) or the marginal ``sum(p, dims)`` (:66-99), 1-based dims.

        The marginal runs on the GPU and returns a device-resident vector whose states are
        in order of first appearance and whose values are summed in state order, as the
        reference's loop sums them.
        """
        if dims is None:
            if self._values is not None:
                return float(np.sum(self._values))
            tot = C.c_double()
            L.call("nc_fspvec_sum", self._dev.handle, C.byref(tot))
            return tot.value
        dims = list(dict.fromkeys(int(d) for d in dims))
        if not dims or not (min(dims) >= 1 and max(dims) <= self._ns):
            raise ArgumentError(L.NC_ERR_ARG, f"Input dimensions must be between 1 and {self._ns}.")
        d = L.i32([k - 1 for k in dims])
        h = C.c_void_p()
        L.call("nc_fspvec_marginal", self._device(), len(d), L.ptr(d, C.c_int32), C.byref(h))
        return FspVectorSparse._from_device(h.value)

    def to_array(self):
        """``Array(p)`` (fspvector.jl:107-129): dense array indexed by molecule counts, built
        on the GPU (column-major, the first species fastest, as Julia's ``Array``)."""
        if self._n == 0:
            raise ArgumentError(L.NC_ERR_ARG, "Cannot construct dense array from empty FspVector instance.")
        bounds = np.zeros(max(self._ns, 1), dtype=np.int64)
        L.call("nc_fspvec_dense", self._device(), L.ptr(bounds, C.c_int64), 0, None)
        shape = tuple(int(b) for b in bounds[: self._ns])
        out = np.empty(int(np.prod(shape, dtype=np.int64)), dtype=np.float64)
        L.call("nc_fspvec_dense", self._device(), L.ptr(bounds, C.c_int64), out.size, L.ptr(out, C.c_double))
        return out.reshape(shape, order="F")

    def __len__(self):
        return self._n

    def __repr__(self):
        return f"Sparse multi-indexed vector with {self.nnz()} nonzero entries."


def nnz(p: FspVectorSparse) -> int:
    return p.nnz()
