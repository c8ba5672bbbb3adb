"""Oracle restatement of NumCME.jl's sparse FSP vector (TEST INFRASTRUCTURE ONLY).

Restates ``src/fspvector/fspvector.jl``:

* ``sum(p)``         ``fspvector.jl:57-59``
* ``sum(p, dims)``   ``fspvector.jl:66-99`` — the reference's loop, line for line: walk the
  states in order, give each new reduced state the next index, start its value with the
  first value and add the later ones in state order.
* ``Array(p)``       ``fspvector.jl:107-129`` — bounds = max count + 1 per species, zeros
  elsewhere, assignments in state order (a repeated state keeps the last value).

Inputs are an ``(n, ns)`` integer state array and ``n`` values; dims are 1-based as in the
reference. Pure Python loops: small inputs only.

Parity pins (``tests/test_oracle_fspvector.py``): the known answers of
``test/test_fspvec.jl`` (sum 11, marginal nnz 11 and 2, the ArgumentErrors), hand-computed
marginals and dense arrays, and, on random inputs, numpy's independent group-by
(``np.unique`` + ``np.add.at``, which also adds in index order) and a brute-force dense
fill.
"""
from __future__ import annotations

import numpy as np


class ArgumentError(ValueError):
    """Julia's ArgumentError."""


def total(values):
    """``sum(p)`` (fspvector.jl:57-59)."""
    return float(np.sum(np.asarray(values, dtype=np.float64)))


def marginal(states, values, dims):
    """``sum(p, dims)`` (fspvector.jl:66-99); returns (reduced states, values)."""
    states = np.asarray(states, dtype=np.int64)
    ns = states.shape[1]
    dims = list(dict.fromkeys(int(d) for d in dims))  # unique(dims), first occurrence order
    if not (min(dims) >= 1 and max(dims) <= ns):
        raise ArgumentError(f"Input dimensions must be between 1 and {ns}.")
    keepdims = [k for k in range(1, ns + 1) if k not in dims]  # setdiff(1:NS, dims)
    reduced_states, reduced_vals, index = [], [], {}
    for i in range(states.shape[0]):
        red = tuple(int(states[i, k - 1]) for k in keepdims)
        idx = index.get(red, 0)
        if idx == 0:
            reduced_states.append(red)
            reduced_vals.append(float(values[i]))
            index[red] = len(reduced_states)
        else:
            reduced_vals[idx - 1] += float(values[i])
    out = np.asarray(reduced_states, dtype=np.int64).reshape(len(reduced_states), len(keepdims))
    return out, np.asarray(reduced_vals, dtype=np.float64)


def dense(states, values):
    """``Array(p)`` (fspvector.jl:107-129), indexed by molecule counts."""
    states = np.asarray(states, dtype=np.int64)
    if states.shape[0] == 0:
        raise ArgumentError("Cannot construct dense array from empty FspVector instance.")
    ns = states.shape[1]
    bounds = [0] * ns
    for x in states:
        for s in range(ns):
            bounds[s] = max(bounds[s], int(x[s]) + 1)
    out = np.zeros(bounds, dtype=np.float64)
    for i in range(states.shape[0]):
        out[tuple(int(c) for c in states[i])] = values[i]
    return out
