"""ℓ1 distance between two sparse FSP vectors over the union of their state sets.

Each vector is given by its states (n × species, non-negative integers, each state once)
and its values; a state present in one vector only contributes the absolute value it has
there. The states are encoded as int64 keys so that a comparison of tens of millions of
states is a sort, not a Python loop.
"""
from __future__ import annotations

import numpy as np


def _keys(states: np.ndarray, base: int) -> np.ndarray:
    k = np.zeros(states.shape[0], dtype=np.int64)
    for c in range(states.shape[1]):
        k = k * base + states[:, c].astype(np.int64)
    return k


def l1_distance(a_states, a_vals, b_states, b_vals) -> float:
    a_states, b_states = np.asarray(a_states), np.asarray(b_states)
    if a_states.ndim == 1:
        a_states = a_states.reshape(-1, 1)
    if b_states.ndim == 1:
        b_states = b_states.reshape(-1, 1)
    assert a_states.shape == (len(a_vals), b_states.shape[1]) and b_states.shape[0] == len(b_vals)
    both = np.concatenate([a_states.ravel(), b_states.ravel()])
    if both.size == 0:
        return 0.0
    assert both.min() >= 0
    base = int(both.max()) + 1
    assert base ** a_states.shape[1] < 2**62, "state coordinates too large for int64 keys"
    keys = np.concatenate([_keys(a_states, base), _keys(b_states, base)])
    vals = np.concatenate([np.asarray(a_vals, dtype=float), -np.asarray(b_vals, dtype=float)])
    _, inv = np.unique(keys, return_inverse=True)
    return float(np.abs(np.bincount(inv.ravel(), weights=vals)).sum())
