"""Oracle of the learned space adapter's inputs (TEST INFRASTRUCTURE ONLY).

The adapter has no reference counterpart. What it restates from the reference:

* the reactions to expand, as ``SelectiveRStepAdapter.adapt!`` picks them
  (rstepadapters.jl:104-105: ``findall(dsinks .> 0)``; an empty list means every reaction,
  sparsestatespace.jl:163), and
* the boundary states, i.e. the first-level explorables of ``expand!`` (states with a sink
  on one of those reactions, in ascending index order, sparsestatespace.jl:165-173).

Features of a boundary state x (p its probability, a_r(t, x) its propensities):
  f0..f2  x_s / (1 + max_s) for species 0..2, max over the whole space (0 when s >= ns)
  f3      log10(max(p, 1e-30)) / 30
  f4      log10(1 + sum_r a_r) / 4
  f5      log10(max(share, 1e-30)) / 30, share = p sum_{r in sinks(x)} a_r / sum_r dsinks_r
  f6      t / tend
  f7      |sinks(x)| / nr
computed in FP64 and rounded to FP32. A record's label is 1 when the probability held by
its successors x + s_r (r a sink reaction of x that was expanded) reaches ``label_mass``.
"""
from __future__ import annotations

import numpy as np

from .fspmatrix import Joint, Separable, TimeInvariant


def expand_reactions(dsinks):
    """Boolean mask of the reactions the selective adapter expands along."""
    sel = np.asarray(dsinks, dtype=np.float64) > 0.0
    if not sel.any():
        sel[:] = True
    return sel


def propensities(props, states, theta, t):
    """(n, nr) propensity values a_r(t, x) of the given states."""
    states = np.asarray(states, dtype=np.int64)
    cols = [states[:, s] for s in range(states.shape[1])]
    n = states.shape[0]
    out = np.empty((n, len(props)))
    for r, pr in enumerate(props):
        if isinstance(pr, TimeInvariant):
            v = pr.f(cols, theta)
        elif isinstance(pr, Separable):
            v = np.asarray(pr.tfactor(t, theta), dtype=np.float64) * np.asarray(pr.statefactor(cols, theta))
        elif isinstance(pr, Joint):
            v = pr.f(t, cols, theta)
        else:
            raise TypeError(pr)
        out[:, r] = np.broadcast_to(np.asarray(v, dtype=np.float64), (n,))
    return out


def boundary(sink_conn, dsinks):
    """Ascending 0-based ids of the boundary states (sink_conn: (n, nr), nonzero = sink)."""
    sel = expand_reactions(dsinks)
    return np.nonzero((np.asarray(sink_conn)[:, sel] != 0).any(axis=1))[0]


def features(states, sink_conn, props, theta, p, dsinks, t, tend):
    states = np.asarray(states, dtype=np.int64)
    sink_conn = np.asarray(sink_conn)
    p = np.asarray(p, dtype=np.float64)
    ids = boundary(sink_conn, dsinks)
    n, ns = states.shape
    nr = sink_conn.shape[1]
    X = states[ids]
    a = propensities(props, X, theta, t)
    sinks = sink_conn[ids] != 0
    F = np.zeros((ids.size, 8))
    xmax = states[:, : min(ns, 3)].max(axis=0) if n else np.zeros(min(ns, 3))
    for s in range(min(ns, 3)):
        F[:, s] = X[:, s] / (1.0 + xmax[s])
    pi = p[ids]
    F[:, 3] = np.log10(np.maximum(pi, 1e-30)) / 30.0
    F[:, 4] = np.log10(1.0 + a.sum(axis=1)) / 4.0
    flux = float(np.sum(dsinks))
    inv = 1.0 / flux if flux > 0 else 0.0
    F[:, 5] = np.log10(np.maximum(pi * (a * sinks).sum(axis=1) * inv, 1e-30)) / 30.0
    F[:, 6] = t / tend
    F[:, 7] = sinks.sum(axis=1) / nr
    return ids, F.astype(np.float32)


def labels(lookup, p_now, stoich, rec_states, rec_masks, label_mass):
    """Labels of records (states (m, ns), masks (m, nr) bool) against a space given by
    ``lookup`` (rows -> 1-based index or 0) and its probabilities ``p_now``."""
    stoich = np.asarray(stoich, dtype=np.int64)
    rec_states = np.asarray(rec_states, dtype=np.int64)
    mass = np.zeros(rec_states.shape[0])
    for r in range(stoich.shape[1]):
        m = np.asarray(rec_masks)[:, r]
        if not m.any():
            continue
        Y = rec_states[m] + stoich[:, r]
        idx = lookup(Y)
        v = np.where(idx > 0, np.asarray(p_now)[np.maximum(idx - 1, 0)], 0.0)
        mass[m] += v
    return (mass >= label_mass).astype(np.float32)
