"""Pins of the oracle's state-space restatement (CPU only).

Sources of truth, none of which is the oracle's own code:
* the reference's state-space tests, test/test_statespace.jl:42-87 (counts, sorted sets);
* hand-derived known-answer index orders for expand! (sparsestatespace.jl:165-191 pops
  explorables from the back of a Deque, so exploration is LIFO), chosen so that a FIFO
  or reaction-order mistake changes the order;
* brute-force reachability: the set after expand!(L) equals the set of non-negative
  states within graph distance L of the initial set along explorable directions;
* the literal and vectorised restatements agree exactly on random call sequences.
"""
import itertools

import numpy as np
import pytest
from hypothesis import given, settings, strategies as st

from oracle.statespace import LiteralStateSpace, StateSpace

BURSTING = np.array([[-1, 1, 0, 0], [1, -1, 0, 0], [0, 0, 1, -1]])
TOGGLE = np.array([[1, -1, 0, 0], [0, 0, 1, -1]])

IMPLS = [LiteralStateSpace, StateSpace]


@pytest.mark.parametrize("cls", IMPLS)
def test_reference_bursting_counts(cls):
    # test_statespace.jl:42-58
    sp = cls(BURSTING, [[1, 0, 0]])
    assert sp.arrays()[0].shape[0] == 1
    sp.expand(0)
    assert sp.arrays()[0].shape[0] == 1
    sp.expand(1)
    assert sp.arrays()[0].shape[0] == 3
    sp = cls(BURSTING, [[1, 0, 0]])
    sp.expand(3)
    assert sp.arrays()[0].shape[0] == 7
    sp = cls(BURSTING, [[1, 0, 0]])
    sp.expand(4)
    expected = sorted([[1, 0, 0], [0, 1, 0], [1, 0, 1], [0, 1, 1], [1, 0, 2], [0, 1, 2], [1, 0, 3], [0, 1, 3], [1, 0, 4]])
    assert sorted(sp.arrays()[0].tolist()) == expected


@pytest.mark.parametrize("cls", IMPLS)
def test_reference_toggle_counts(cls):
    # test_statespace.jl:61-87
    sp = cls(TOGGLE, [[0, 0]])
    sp.expand(0)
    assert sp.arrays()[0].shape[0] == 1
    sp.expand(1)
    assert sp.arrays()[0].shape[0] == 3
    sp = cls(TOGGLE, [[0, 0]])
    sp.expand(3)
    assert sp.arrays()[0].shape[0] == 10
    expected = sorted([[0, 0], [1, 0], [2, 0], [3, 0], [0, 1], [1, 1], [2, 1], [0, 2], [1, 2], [0, 3]])
    assert sorted(sp.arrays()[0].tolist()) == expected


# Known answers derived by hand from the code text (LIFO explorables, reactions in order):
# level 1 from (0,0): (1,0),(0,1); level 2 pops (0,1) first: (1,1),(0,2), then (1,0): (2,0);
# level 3 pops (2,0): (3,0),(2,1); (0,2): (1,2),(0,3); (1,1): nothing new.
TOGGLE_L3_ORDER = [[0, 0], [1, 0], [0, 1], [1, 1], [0, 2], [2, 0], [3, 0], [2, 1], [1, 2], [0, 3]]
BURSTING_L4_ORDER = [[1, 0, 0], [0, 1, 0], [1, 0, 1], [0, 1, 1], [1, 0, 2], [0, 1, 2], [1, 0, 3], [0, 1, 3], [1, 0, 4]]


@pytest.mark.parametrize("cls", IMPLS)
def test_known_answer_orders(cls):
    sp = cls(TOGGLE, [[0, 0]])
    sp.expand(3)
    assert sp.arrays()[0].tolist() == TOGGLE_L3_ORDER
    sp = cls(BURSTING, [[1, 0, 0]])
    sp.expand(4)
    assert sp.arrays()[0].tolist() == BURSTING_L4_ORDER


@pytest.mark.parametrize("cls", IMPLS)
def test_known_answer_only_reactions(cls):
    # onlyreactions = [1, 3] (births): from (0,0), level 1 -> (1,0),(0,1); level 2 pops (0,1): (1,1),(0,2); then (1,0): (2,0)
    sp = cls(TOGGLE, [[0, 0]])
    sp.expand(2, onlyreactions=[1, 3])
    assert sp.arrays()[0].tolist() == [[0, 0], [1, 0], [0, 1], [1, 1], [0, 2], [2, 0]]


@pytest.mark.parametrize("cls", IMPLS)
def test_known_answer_connectivity(cls):
    sp = cls(TOGGLE, [[0, 0]])
    sp.expand(1)
    states, sc, sk = sp.arrays()
    # states (0,0)=1, (1,0)=2, (0,1)=3
    # state_conn[i][r] = index of x_i - s_r
    assert sc.tolist() == [[0, 2, 0, 3], [1, 0, 0, 0], [0, 0, 1, 0]]
    # sinks: (1,0)+s1=(2,0) absent, (1,0)+s3=(1,1) absent; (0,1)+s1=(1,1), (0,1)+s3=(0,2) absent
    assert sk.tolist() == [[0, 0, 0, 0], [1, 0, 3, 0], [1, 0, 3, 0]]


@pytest.mark.parametrize("cls", IMPLS)
def test_known_answer_delete(cls):
    sp = cls(TOGGLE, [[0, 0]])
    sp.expand(1)
    sp.deleteat([2])  # remove (1,0)
    states, sc, sk = sp.arrays()
    assert states.tolist() == [[0, 0], [0, 1]]
    # (0,0): predecessor via r2 was (1,0) (deleted) -> 0; successor via r1 now absent -> sink 1
    assert sc.tolist() == [[0, 0, 0, 2], [0, 0, 1, 0]]
    assert sk.tolist() == [[1, 0, 0, 0], [1, 0, 3, 0]]


def _brute_reach(S, init, levels, rs):
    cur = {tuple(x) for x in init}
    frontier = set(cur)
    for _ in range(levels):
        nxt = set()
        for x in frontier:
            for r in rs:
                y = tuple(int(a + b) for a, b in zip(x, S[:, r]))
                if min(y) >= 0 and y not in cur:
                    nxt.add(y)
        cur |= nxt
        frontier = nxt
    return cur


@pytest.mark.parametrize("levels", [1, 2, 5, 9])
def test_brute_force_reachability(levels):
    for S, x0 in ((TOGGLE, [0, 0]), (BURSTING, [1, 0, 0]), (TOGGLE, [3, 2])):
        sp = StateSpace(S, [x0])
        sp.expand(levels)
        got = {tuple(x) for x in sp.arrays()[0].tolist()}
        assert got == _brute_reach(S, [x0], levels, range(S.shape[1]))


def _cross_check(S, init, ops):
    a, b = LiteralStateSpace(S, init), StateSpace(S, init)
    for op in ops:
        if op[0] == "expand":
            a.expand(op[1], op[2]), b.expand(op[1], op[2])
        else:
            n = len(a.states)
            ids = sorted({i for i in op[1] if 1 <= i <= n})
            a.deleteat(ids), b.deleteat(ids)
        for u, v in zip(a.arrays(), b.arrays()):
            assert np.array_equal(u, v)


op_strategy = st.one_of(
    st.tuples(st.just("expand"), st.integers(0, 3), st.lists(st.integers(1, 4), max_size=3, unique=True)),
    st.tuples(st.just("delete"), st.lists(st.integers(1, 40), max_size=6)),
)


@settings(max_examples=60, deadline=None)
@given(ops=st.lists(op_strategy, min_size=1, max_size=6),
       init=st.lists(st.tuples(st.integers(-1, 4), st.integers(-1, 4)), min_size=1, max_size=5))
def test_literal_vs_vectorised_toggle(ops, init):
    _cross_check(TOGGLE, [list(x) for x in init], ops)


@settings(max_examples=30, deadline=None)
@given(ops=st.lists(op_strategy, min_size=1, max_size=5))
def test_literal_vs_vectorised_bursting(ops):
    _cross_check(BURSTING, [[1, 0, 0]], ops)


def test_duplicates_and_negatives_dropped():
    # _addstates! keeps the first occurrence and drops states with a negative entry
    for cls in IMPLS:
        sp = cls(TOGGLE, [[2, 1], [0, -1], [2, 1], [0, 0]])
        assert sp.arrays()[0].tolist() == [[2, 1], [0, 0]]
