"""Sparse multi-indexed probability vectors — mirrors ``src/fspvector/fspvector.jl``.

``FspVectorSparse`` is the snapshot type ``solve`` returns: a state list (rows of an
int64 array), values, and a state -> index map built on demand.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from ._lib import ArgumentError


class FspVectorSparse:
    """``FspVectorSparse(states, values; checksizes=true)`` (fspvector.jl:13-35)."""

    def __init__(self, states, values, checksizes=True, copy=True):
        st = np.array(states, dtype=np.int64, copy=copy)
        if st.ndim == 1:
            st = st.reshape(len(values) if len(values) else 0, -1) if st.size else st.reshape(0, 0)
        vals = np.array(values, dtype=np.float64, copy=copy).ravel()
        if checksizes and st.shape[0] != vals.shape[0]:
            raise ArgumentError(L.NC_ERR_ARG, "State and value lists must have equal lengths.")
        self.states = st
        self.values = vals
        self._dict = None

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
        return self.states.shape[0]

    def sum(self, dims=None):
        """``sum(p)`` (fspvector.jl:57-59) or the marginal ``sum(p, dims)`` (:66-99), 1-based dims."""
        if dims is None:
            return float(np.sum(self.values))
        ns = self.states.shape[1]
        dims = list(dict.fromkeys(int(d) for d in dims))
        if not (min(dims) >= 1 and max(dims) <= ns):
            raise ArgumentError(L.NC_ERR_ARG, f"Input dimensions must be between 1 and {ns}.")
        keep = [k for k in range(ns) if (k + 1) not in dims]
        red = self.states[:, keep]
        if red.shape[0] == 0:
            return FspVectorSparse(red, np.zeros(0), copy=False)
        # first-appearance order of the reduced states, as the reference's loop
        _, first, inv = np.unique(red, axis=0, return_index=True, return_inverse=True)
        inv = inv.ravel()
        order = np.argsort(first, kind="stable")
        rank = np.empty_like(order)
        rank[order] = np.arange(order.size)
        vals = np.zeros(order.size)
        np.add.at(vals, rank[inv], self.values)
        return FspVectorSparse(red[np.sort(first)], vals, copy=False)

    def to_array(self):
        """``Array(p)`` (fspvector.jl:107-129): dense array indexed by molecule counts."""
        if self.states.shape[0] == 0:
            raise ArgumentError(L.NC_ERR_ARG, "Cannot construct dense array from empty FspVector instance.")
        bounds = self.states.max(axis=0) + 1
        out = np.zeros(tuple(bounds), dtype=np.float64)
        out[tuple(self.states.T)] = self.values
        return out

    def __len__(self):
        return self.states.shape[0]

    def __repr__(self):
        return f"Sparse multi-indexed vector with {self.nnz()} nonzero entries."


def nnz(p: FspVectorSparse) -> int:
    return p.nnz()
