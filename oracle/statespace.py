"""Oracle restatement of NumCME.jl's sparse FSP state space (TEST INFRASTRUCTURE ONLY).

Restates ``src/statespace/sparse/sparsestatespace.jl``:

* constructor           ``sparsestatespace.jl:103-124`` (single state: ``:142-144``)
* ``expand!``           ``sparsestatespace.jl:153-194``
* ``_is_all_nonnegative`` ``sparsestatespace.jl:197-202``
* ``_addstates!``       ``sparsestatespace.jl:208-267``
* ``deleteat!``         ``sparsestatespace.jl:276-331``

Conventions follow the reference: indices are 1-based, ``state_conn[i][r]`` is
the 1-based index of ``x_i - s_r`` or 0, ``sink_conn[i][r]`` is ``r+1`` (the
1-based reaction number) when ``x_i + s_r`` is non-negative and not in the
space, else 0.

Two implementations with identical semantics:

``LiteralStateSpace``
    a line-by-line pure-Python restatement (dict + deque), for small inputs.
``StateSpace``
    a vectorised numpy restatement for large inputs. First-occurrence
    deduplication uses ``np.unique(..., return_index=True)``; connectivity is
    filled with whole-array operations, which is valid because the reference's
    connectivity loop (``:241-266``) writes values that depend only on the
    final state dictionary, not on the loop order.

Parity pins (see ``tests/test_oracle_statespace.py``): the reference's own
state-space tests (``test/test_statespace.jl:42-87``: counts 1/3/7 and the
level-4 telegraph set; counts 1/3/10 and the level-3 toggle set), hand-derived
known-answer orderings that separate LIFO from FIFO exploration, brute-force
reachability sets on small lattices, and cross-checks of the two
implementations under random call sequences.
"""
from __future__ import annotations

from collections import deque

import numpy as np


def _is_all_nonnegative(x) -> bool:
    # sparsestatespace.jl:197-202
    for xi in x:
        if xi < 0:
            return False
    return True


class LiteralStateSpace:
    """Line-by-line restatement of ``StateSpaceSparse`` (pure Python, small inputs)."""

    def __init__(self, stoich, init_states):
        # sparsestatespace.jl:103-124; ``stoich`` is NS x NR (column r = s_r).
        self.stoich = [list(map(int, row)) for row in np.asarray(stoich)]
        self.ns = len(self.stoich)
        self.nr = len(self.stoich[0])
        self.sink_count = self.nr
        self.states: list[tuple] = []
        self.state2idx: dict[tuple, int] = {}
        self.state_conn: list[list[int]] = []
        self.sink_conn: list[list[int]] = []
        self._addstates([tuple(int(v) for v in s) for s in init_states])

    def col(self, r):
        return [self.stoich[s][r] for s in range(self.ns)]

    def _addstates(self, newstates):
        # sparsestatespace.jl:208-267
        unique_new = []
        newidx = len(self.states)
        for state in newstates:
            if self.state2idx.get(state, 0) == 0 and _is_all_nonnegative(state):
                newidx += 1
                unique_new.append(state)
                self.state2idx[state] = newidx
        old = len(self.states)
        self.states.extend(unique_new)
        for _ in unique_new:
            self.state_conn.append([0] * self.nr)
            self.sink_conn.append([0] * self.nr)
        for ni in range(old + 1, len(self.states) + 1):
            x = self.states[ni - 1]
            for ir in range(self.nr):
                s = self.col(ir)
                y = tuple(a - b for a, b in zip(x, s))
                ridx = self.state2idx.get(y, 0)
                if ridx != 0:
                    self.state_conn[ni - 1][ir] = ridx
                    self.sink_conn[ridx - 1][ir] = 0
            for ir in range(self.nr):
                s = self.col(ir)
                z = tuple(a + b for a, b in zip(x, s))
                if _is_all_nonnegative(z):
                    ridx = self.state2idx.get(z, 0)
                    if ridx == 0:
                        self.sink_conn[ni - 1][ir] = ir + 1
                    else:
                        self.state_conn[ridx - 1][ir] = ni

    def expand(self, levels, onlyreactions=()):
        # sparsestatespace.jl:153-194; ``onlyreactions`` are 1-based like the reference.
        if levels <= 0:
            return
        rs = [r - 1 for r in onlyreactions] if len(onlyreactions) else list(range(self.nr))
        explorables = deque()
        for idx in range(1, len(self.states) + 1):
            for ir in rs:
                if self.sink_conn[idx - 1][ir] != 0:
                    explorables.append(idx)
                    break
        for _level in range(levels):
            cands = []
            while explorables:
                idx = explorables.pop()  # Deque pop! takes from the back (LIFO)
                for ir in rs:
                    cands.append(tuple(a + b for a, b in zip(self.states[idx - 1], self.col(ir))))
            old = len(self.states)
            self._addstates(cands)
            for idx in range(old + 1, len(self.states) + 1):
                explorables.append(idx)

    def deleteat(self, ids):
        # sparsestatespace.jl:276-331; ``ids`` 1-based, must be sorted after unique
        ids = list(dict.fromkeys(int(i) for i in ids))
        if ids != sorted(ids):
            raise ValueError("indices must be unique and sorted")
        old = len(self.states)
        newidxs = list(range(1, old + 1))
        for i in ids:
            newidxs[i - 1] = 0
        idx = 1
        for i in range(old):
            if newidxs[i] != 0:
                newidxs[i] = idx
                idx += 1
        for i in ids:
            del self.state2idx[self.states[i - 1]]
        dead = set(ids)
        self.states = [s for k, s in enumerate(self.states, 1) if k not in dead]
        self.state_conn = [c for k, c in enumerate(self.state_conn, 1) if k not in dead]
        self.sink_conn = [c for k, c in enumerate(self.sink_conn, 1) if k not in dead]
        if not self.states:
            return
        for i, x in enumerate(self.states):
            self.state2idx[x] = newidxs[self.state2idx[x] - 1]
            for ir in range(self.nr):
                o = self.state_conn[i][ir]
                self.state_conn[i][ir] = newidxs[o - 1] if o != 0 else 0
        for i, x in enumerate(self.states):
            for ir in range(self.nr):
                if self.sink_conn[i][ir] == 0:
                    z = tuple(a + b for a, b in zip(x, self.col(ir)))
                    if _is_all_nonnegative(z) and self.state2idx.get(z, 0) == 0:
                        self.sink_conn[i][ir] = ir + 1

    # array views, same layout as ``StateSpace``
    def arrays(self):
        return (np.array(self.states, dtype=np.int64).reshape(-1, self.ns),
                np.array(self.state_conn, dtype=np.int64).reshape(-1, self.nr),
                np.array(self.sink_conn, dtype=np.int64).reshape(-1, self.nr))


class StateSpace:
    """Vectorised numpy restatement of ``StateSpaceSparse`` (same semantics)."""

    def __init__(self, stoich, init_states):
        self.stoich = np.asarray(stoich, dtype=np.int64)  # NS x NR
        self.ns, self.nr = self.stoich.shape
        self.sink_count = self.nr
        self.bits = 64 // self.ns if self.ns > 1 else 62
        self.states = np.zeros((0, self.ns), dtype=np.int64)
        self.state_conn = np.zeros((0, self.nr), dtype=np.int64)
        self.sink_conn = np.zeros((0, self.nr), dtype=np.int64)
        self._keys_sorted = np.zeros(0, dtype=np.int64)
        self._order = np.zeros(0, dtype=np.int64)
        init = np.asarray(init_states, dtype=np.int64).reshape(-1, self.ns)
        self._addstates(init)

    # --- lookup helpers -------------------------------------------------------------
    def _pack(self, X):
        if X.size and X.max() >= (1 << (self.bits - 1)):
            raise OverflowError("state entry too large for the oracle's key packing")
        k = np.zeros(X.shape[0], dtype=np.int64)
        for s in range(self.ns):
            k = (k << self.bits) | X[:, s]
        return k

    def _reindex(self):
        keys = self._pack(self.states)
        self._order = np.argsort(keys, kind="stable")
        self._keys_sorted = keys[self._order]

    def lookup(self, X):
        """1-based index of each row of ``X`` in the space, 0 if absent or negative."""
        X = np.asarray(X, dtype=np.int64).reshape(-1, self.ns)
        out = np.zeros(X.shape[0], dtype=np.int64)
        # states beyond the packing range cannot have been added (adding them raises)
        ok = np.all(X >= 0, axis=1) & np.all(X < (1 << (self.bits - 1)), axis=1)
        if not ok.any() or self._keys_sorted.size == 0:
            return out
        k = self._pack(X[ok])
        pos = np.searchsorted(self._keys_sorted, k)
        pos_c = np.minimum(pos, self._keys_sorted.size - 1)
        hit = self._keys_sorted[pos_c] == k
        res = np.where(hit, self._order[pos_c] + 1, 0)
        out[ok] = res
        return out

    @property
    def count(self):
        return self.states.shape[0]

    # --- _addstates! ------------------------------------------------------------------
    def _addstates(self, cands):
        # sparsestatespace.jl:208-267
        cands = np.asarray(cands, dtype=np.int64).reshape(-1, self.ns)
        nonneg = np.all(cands >= 0, axis=1)
        absent = self.lookup(cands) == 0
        pos = np.nonzero(nonneg & absent)[0]
        if pos.size:
            keys = self._pack(cands[pos])
            _, first = np.unique(keys, return_index=True)
            pos = pos[np.sort(first)]  # first occurrence wins, candidate order kept
        new = cands[pos]
        old = self.count
        n_new = new.shape[0]
        self.states = np.vstack([self.states, new])
        self.state_conn = np.vstack([self.state_conn, np.zeros((n_new, self.nr), np.int64)])
        self.sink_conn = np.vstack([self.sink_conn, np.zeros((n_new, self.nr), np.int64)])
        self._reindex()
        if n_new == 0:
            return
        new_ids = np.arange(old + 1, old + n_new + 1)
        for ir in range(self.nr):
            s = self.stoich[:, ir]
            ridx = self.lookup(new - s)
            hit = ridx != 0
            self.state_conn[new_ids[hit] - 1, ir] = ridx[hit]
            self.sink_conn[ridx[hit] - 1, ir] = 0
        for ir in range(self.nr):
            s = self.stoich[:, ir]
            z = new + s
            nonneg_z = np.all(z >= 0, axis=1)
            ridx = self.lookup(z)
            sink = nonneg_z & (ridx == 0)
            self.sink_conn[new_ids[sink] - 1, ir] = ir + 1
            hit = ridx != 0
            self.state_conn[ridx[hit] - 1, ir] = new_ids[hit]

    # --- expand! ----------------------------------------------------------------------
    def expand(self, levels, onlyreactions=()):
        # sparsestatespace.jl:153-194; ``onlyreactions`` 1-based
        if levels <= 0:
            return
        rs = np.asarray([r - 1 for r in onlyreactions] if len(onlyreactions) else range(self.nr), dtype=np.int64)
        explorables = np.nonzero(np.any(self.sink_conn[:, rs] != 0, axis=1))[0] + 1
        for _level in range(levels):
            E = explorables[::-1]  # LIFO: last pushed explorable is popped first
            cands = (self.states[E - 1][:, None, :] + self.stoich[:, rs].T[None, :, :]).reshape(-1, self.ns)
            old = self.count
            self._addstates(cands)
            explorables = np.arange(old + 1, self.count + 1)

    # --- deleteat! --------------------------------------------------------------------
    def deleteat(self, ids):
        # sparsestatespace.jl:276-331; ``ids`` 1-based
        ids = np.asarray(ids, dtype=np.int64)
        _, first = np.unique(ids, return_index=True)
        ids = ids[np.sort(first)]
        if ids.size and np.any(np.diff(ids) < 0):
            raise ValueError("indices must be unique and sorted")
        old = self.count
        keep = np.ones(old, dtype=bool)
        keep[ids - 1] = False
        newidxs = np.zeros(old + 1, dtype=np.int64)  # slot 0 maps 0 -> 0
        newidxs[1:][keep] = np.arange(1, keep.sum() + 1)
        self.states = self.states[keep]
        self.state_conn = newidxs[self.state_conn[keep]]
        self.sink_conn = self.sink_conn[keep]
        self._reindex()
        if self.count == 0:
            return
        for ir in range(self.nr):
            z = self.states + self.stoich[:, ir]
            cand = (self.sink_conn[:, ir] == 0) & np.all(z >= 0, axis=1) & (self.lookup(z) == 0)
            self.sink_conn[cand, ir] = ir + 1

    def arrays(self):
        return self.states.copy(), self.state_conn.copy(), self.sink_conn.copy()


def rectangular_states(bounds, order="first_fastest"):
    """All states of the box ``0 <= x_s <= bounds[s]`` with species 0 varying fastest."""
    grids = np.meshgrid(*[np.arange(b + 1) for b in bounds[::-1]], indexing="ij")
    X = np.stack([g.ravel() for g in grids[::-1]], axis=1)
    return X.astype(np.int64)
