"""Pins of the learned-adapter oracle (CPU): the boundary choice against the reference's
selective expansion rule, and the features on a hand-computed birth-death case."""
import numpy as np

import models
from oracle import learned as ol
from oracle.fspmatrix import TimeInvariant
from oracle.statespace import StateSpace


def test_expand_reactions_rule():
    # findall(dsinks .> 0); empty -> every reaction
    np.testing.assert_array_equal(ol.expand_reactions([0.0, 1e-9, -1e-3]), [False, True, False])
    np.testing.assert_array_equal(ol.expand_reactions([0.0, -0.0, 0.0]), [True, True, True])


def test_boundary_is_first_level_explorables():
    # the states expand! would explore first are exactly the oracle's boundary states
    S, props, th = models.cascade()
    sp = StateSpace(S, [[0, 0, 0]])
    sp.expand(4)
    states, _, sink_conn = sp.states, sp.state_conn, sp.sink_conn
    ds = np.array([1.0, 0.0, 2.0, 0.0, 0.0, 0.0])
    ids = ol.boundary(sink_conn, ds)
    manual = [i for i in range(len(states)) if sink_conn[i, 0] or sink_conn[i, 2]]
    np.testing.assert_array_equal(ids, manual)
    # after expanding along reactions 0 and 2 every such state gains its successors
    before = {tuple(x) for x in states}
    sp.expand(1, onlyreactions=[1, 3])  # 1-based reaction ids of the reference API
    after = {tuple(x) for x in sp.states}
    for i in ids:
        for r in (0, 2):
            y = tuple(states[i] + S[:, r])
            if sink_conn[i, r]:
                assert y not in before and y in after


def test_features_known_answer():
    # birth-death 0 -> X (k = 2), X -> 0 (g = 0.5 x) on states 0..4: only x = 4 has a sink
    S = np.array([[1, -1]])
    props = [TimeInvariant(lambda x, p: p[0] + 0.0 * x[0]), TimeInvariant(lambda x, p: p[1] * x[0])]
    th = np.array([2.0, 0.5])
    states = np.arange(5).reshape(-1, 1)
    sink_conn = np.zeros((5, 2), dtype=int)
    sink_conn[4, 0] = 1
    p = np.array([0.1, 0.2, 0.3, 0.3, 0.1])
    dsinks = np.array([0.1 * 2.0, 0.0])
    ids, F = ol.features(states, sink_conn, props, th, p, dsinks, t=3.0, tend=12.0)
    assert ids.tolist() == [4]
    expect = [4 / 5, 0, 0, np.log10(0.1) / 30, np.log10(1 + 2.0 + 2.0) / 4, 0.0, 0.25, 0.5]
    np.testing.assert_allclose(F[0], expect, rtol=1e-6, atol=1e-7)


def test_labels_known_answer():
    S = np.array([[1, -1]])
    sp = StateSpace(S, [[0], [1], [2], [3]])
    p = np.array([0.5, 0.2, 1e-3, 1e-12])
    rec = np.array([[1], [2], [3]])
    masks = np.array([[True, False], [True, False], [True, True]])
    y = ol.labels(sp.lookup, p, S, rec, masks, label_mass=1e-6)
    # successor of 1 is 2 (1e-3), of 2 is 3 (1e-12), of 3: 4 (absent) and 2 (1e-3)
    np.testing.assert_array_equal(y, [1, 0, 1])
