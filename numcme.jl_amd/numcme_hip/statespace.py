"""Sparse FSP state space — mirrors ``src/statespace/sparse/sparsestatespace.jl``.

The states, the state -> index map and the connectivity live on the GPU
(``nc_space_*``); this class keeps the reference's API. Host views follow the
reference's 1-based conventions: ``get_state_connectivity(space)[i, r]`` is the 1-based
index of ``x_i - s_r`` or 0, ``get_sink_connectivity(space)[i, r]`` is ``r+1`` or 0, and
``expand(space, levels, onlyreactions)`` / ``deleteat(space, ids)`` take 1-based
reaction numbers and state indices, like ``expand!`` / ``deleteat!``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import ArgumentError


class StateSpaceSparse(L.Handle):
    """``StateSpaceSparse(stoich_matrix, initstates)`` (sparsestatespace.jl:103-124,142-144)."""

    _destroy = "nc_space_destroy"

    def __init__(self, stoich_matrix, initstates):
        S = np.asarray(stoich_matrix, dtype=np.int64)
        if S.ndim != 2:
            raise ArgumentError(L.NC_ERR_ARG, "stoichiometry matrix must be 2-D (species x reactions)")
        ns, nr = S.shape
        X = np.asarray(initstates, dtype=np.int64)
        if X.ndim == 1 and X.size == ns and not isinstance(initstates[0], (list, tuple, np.ndarray)):
            X = X.reshape(1, ns)  # single-state constructor (:142-144)
        X = X.reshape(-1, ns)
        # min and max rather than abs: no temporary the size of the state list
        if (X.max(initial=0) >= 2**31 or X.min(initial=0) <= -2**31
                or S.max(initial=0) >= 2**31 or S.min(initial=0) <= -2**31):
            raise ArgumentError(L.NC_ERR_ARG, "state entries must fit in int32")
        self.stoich_matrix = S
        self.species_count, self.reaction_count = ns, nr
        Sc = L.i32(S.T.copy())
        Xc = L.i32(X)
        h = C.c_void_p()
        L.call("nc_space_create", ns, nr, L.ptr(Sc, C.c_int32), Xc.shape[0], L.ptr(Xc, C.c_int32), C.byref(h))
        super().__init__(h.value)

    # -- reference getters (sparsestatespace.jl:48-80) ------------------------------------
    def get_stoich_matrix(self):
        return self.stoich_matrix

    def get_state_count(self) -> int:
        n = C.c_int64()
        L.call("nc_space_count", self.handle, C.byref(n), None)
        return n.value

    def get_sink_count(self) -> int:
        return self.reaction_count

    def get_states(self) -> np.ndarray:
        n = self.get_state_count()
        out = np.empty((n, self.species_count), dtype=np.int32)
        L.call("nc_space_get_states", self.handle, L.ptr(out, C.c_int32))
        return out.astype(np.int64)

    @property
    def states(self):
        return self.get_states()

    def get_state_connectivity(self) -> np.ndarray:
        n = self.get_state_count()
        out = np.empty((n, self.reaction_count), dtype=np.int32)
        L.call("nc_space_get_state_conn", self.handle, L.ptr(out, C.c_int32))
        return out.astype(np.int64) + 1  # -1 (none) -> 0, index -> 1-based

    def get_sink_connectivity(self) -> np.ndarray:
        n = self.get_state_count()
        out = np.empty((n, self.reaction_count), dtype=np.int32)
        L.call("nc_space_get_sink_conn", self.handle, L.ptr(out, C.c_int32))
        return out.astype(np.int64) + 1  # -1 -> 0, r -> r+1

    def lookup(self, states) -> np.ndarray:
        """1-based index of each state, 0 if absent (``get(state2idx, x, 0)``)."""
        X = L.i32(np.asarray(states).reshape(-1, self.species_count))
        out = np.empty(X.shape[0], dtype=np.int64)
        L.call("nc_space_lookup", self.handle, X.shape[0], L.ptr(X, C.c_int32), L.ptr(out, C.c_int64))
        return out + 1

    def get_statedict(self):
        """The state -> 1-based index map as a Python dict (built on the host; O(n))."""
        return {tuple(x): i + 1 for i, x in enumerate(self.get_states().tolist())}

    # -- mutation -------------------------------------------------------------------------
    def expand(self, expansionlevel: int, onlyreactions=()):
        """``expand!`` (sparsestatespace.jl:153-194); ``onlyreactions`` 1-based."""
        rl = [int(r) - 1 for r in onlyreactions]
        for r in rl:
            if not 0 <= r < self.reaction_count:
                raise ArgumentError(L.NC_ERR_ARG, f"reaction {r + 1} out of range")
        arr = L.i32(rl if rl else [0])
        L.call("nc_space_expand", self.handle, int(expansionlevel), len(rl), L.ptr(arr, C.c_int32))

    def deleteat(self, ids):
        """``deleteat!`` (sparsestatespace.jl:276-331); ``ids`` 1-based, sorted after ``unique``."""
        ids = np.asarray(ids, dtype=np.int64).ravel()
        _, first = np.unique(ids, return_index=True)
        ids = ids[np.sort(first)]
        if ids.size and np.any(np.diff(ids) <= 0):
            raise ArgumentError(L.NC_ERR_ARG, "indices must be unique and sorted")
        z = L.i64(ids - 1 if ids.size else np.zeros(1))
        L.call("nc_space_delete", self.handle, ids.size, L.ptr(z, C.c_int64))


def get_state_count(space): return space.get_state_count()
def get_sink_count(space): return space.get_sink_count()
def get_states(space): return space.get_states()
def get_stoich_matrix(space): return space.get_stoich_matrix()
def get_state_connectivity(space): return space.get_state_connectivity()
def get_sink_connectivity(space): return space.get_sink_connectivity()
def get_statedict(space): return space.get_statedict()
def expand(space, expansionlevel, onlyreactions=()): return space.expand(expansionlevel, onlyreactions)
def deleteat(space, ids): return space.deleteat(ids)
