"""Pins of the oracle's adaptive FSP restatement (CPU only).

References: src/transientcme/sparse/rstepadapters.jl:23-110 (drop rule, init!/adapt!),
src/transientcme/sparse/fspsolve.jl:105-197 (driver), test/test_solver.jl:79-100.
"""
import numpy as np
import pytest
from hypothesis import given, settings, strategies as st

from models import box_states, telegraph, telegraph_tv, to_oracle
from oracle import adaptive as OA
from oracle.solve import solve_fixed
from oracle.statespace import LiteralStateSpace, StateSpace


def literal_dropcount(p, t, tend, fsptol, strict):
    """The reference expression with numpy's sequential sum/cumsum."""
    perm = np.argsort(OA.order_keys(p), kind="stable")
    rest = np.sum(p) - np.cumsum(p[perm])
    thr = 1.0 - t * fsptol / tend
    return perm, int(np.count_nonzero(rest > thr if strict else rest >= thr)), rest, thr


def test_drop_rule_hand_example_separates_ge_from_gt():
    # sorted: 0.0 (2), 0.1 (1), 0.4 (3), 0.5 (0); sum - cumsum = 1.0, 0.9, 0.5, 0.0;
    # with t*fsptol/tend = 0.1 the threshold is 1 - 0.1 == 1.0 - 0.1 exactly, so the
    # second entry ties: RStep (>=) drops two states, SelectiveRStep (>) drops one
    p = np.array([0.5, 0.1, 0.0, 0.4])
    assert OA.drop_ids(p, 1.0, 1.0, 0.1, strict=False).tolist() == [1, 2]
    assert OA.drop_ids(p, 1.0, 1.0, 0.1, strict=True).tolist() == [2]
    # thr = 1 - 0.6 = 0.4: sum - cumsum >= 0.4 for the first three sorted entries
    assert OA.drop_ids(p, 1.0, 1.0, 0.6, strict=False).tolist() == [1, 2, 3]
    assert OA.drop_ids(p, 0.0, 1.0, 0.1, strict=False).tolist() == [2]  # thr = 1: only the zero goes


def test_drop_rule_negative_zero_and_ties_follow_isless():
    p = np.array([0.0, -0.0, 0.25, -1e-20, 0.25, 0.5])
    perm = np.argsort(OA.order_keys(p), kind="stable")
    # Julia isless: -1e-20 < -0.0 < 0.0 < 0.25 (index order kept for ties) < 0.5
    assert perm.tolist() == [3, 1, 0, 2, 4, 5]


@settings(max_examples=60, deadline=None)
@given(st.integers(1, 3000), st.integers(0, 2**32 - 1), st.floats(0.0, 1.0), st.booleans())
def test_drop_rule_matches_literal_formula_away_from_ties(n, seed, frac, strict):
    rng = np.random.default_rng(seed)
    p = rng.exponential(size=n) ** 3
    p[rng.random(n) < 0.2] = 0.0
    if p.sum() == 0:
        p[0] = 1.0
    p /= p.sum()
    tend, t = 10.0, 10.0 * frac
    fsptol = 0.5
    perm, cnt, rest, thr = literal_dropcount(p, t, tend, fsptol, strict)
    # skip draws where a comparison sits within rounding of the threshold
    if np.any(np.abs(rest - thr) < 1e-12):
        return
    assert OA.drop_ids(p, t, tend, fsptol, strict).tolist() == np.sort(perm[:cnt]).tolist()


@pytest.mark.parametrize("n", [1, 7, 1024, 1025, 4096 + 3, 1024 * 1024 + 17])
def test_fixed_scan_exact_on_integers(n):
    # integer-valued doubles below 2^53 add exactly in any order, so every association
    # order must give numpy's sequential cumsum bit for bit
    rng = np.random.default_rng(n)
    x = rng.integers(0, 1000, size=n).astype(np.float64)
    assert np.array_equal(OA.fixed_scan(x), np.cumsum(x))


def test_fixed_scan_close_on_reals():
    rng = np.random.default_rng(3)
    x = rng.random(3 * 1024 * 1024 + 5)
    np.testing.assert_allclose(OA.fixed_scan(x), np.cumsum(x), rtol=1e-12)


def test_adapt_expands_selected_reactions_only():
    S, _, _ = telegraph()
    sp = StateSpace(S, [[1, 0, 0]])
    sp.expand(2)
    lit = LiteralStateSpace(S, [[1, 0, 0]])
    lit.expand(2)
    p = np.full(sp.count, 1.0 / sp.count)
    dsinks = np.array([0.0, 1.0, 0.0, 0.0])  # only reaction 2 (G_on -> G_off) grows
    p_new, ids = OA.adapt(sp, OA.SELECTIVE, 3, False, p, 1.0, 10.0, 1e-4, dsinks)
    lit.expand(3, onlyreactions=(2,))
    assert np.array_equal(sp.states, np.asarray(lit.states))
    assert p_new.size == sp.count and np.all(p_new[p.size:] == 0) and ids.size == 0


def test_adapt_with_no_growing_sink_expands_all():
    S, _, _ = telegraph()
    sp = StateSpace(S, [[1, 0, 0]])
    lit = LiteralStateSpace(S, [[1, 0, 0]])
    OA.adapt(sp, OA.SELECTIVE, 2, False, np.array([1.0]), 1.0, 10.0, 1e-4, np.zeros(4))
    lit.expand(2)
    assert np.array_equal(sp.states, np.asarray(lit.states))


def test_adapt_selective_rejects_empty_p():
    S, _, _ = telegraph()
    sp = StateSpace(S, [[1, 0, 0]])
    sp.deleteat([1])
    with pytest.raises(ValueError):
        OA.adapt(sp, OA.SELECTIVE, 1, True, np.zeros(0), 1.0, 2.0, 1e-4, np.zeros(4))


def test_restart_sinks_eps_trick():
    s = OA.restart_sinks([1e-5, 0.0], 1.0, 1.0, 1e-5)  # sum >= budget -> shifted by eps
    assert s.tolist() == [1e-5 - OA.EPS, -OA.EPS]
    s = OA.restart_sinks([1e-6, 0.0], 1.0, 1.0, 1e-5)
    assert s.tolist() == [1e-6, 0.0]


@pytest.mark.parametrize("kind", [OA.RSTEP, OA.SELECTIVE])
@pytest.mark.parametrize("drop", [False, True])
def test_solve_adaptive_invariants(kind, drop):
    # test/test_solver.jl:79-96: mass conservation on every snapshot; the sink budget; the
    # duplicated final snapshot (fspsolve.jl:163-167 plus :189-194)
    S, props, th = telegraph((0.05, 0.1, 5.0, 1.0))
    ts = np.arange(0.0, 121.0, 20.0)
    fsptol = 1e-4
    out = OA.solve_adaptive(S, to_oracle(props), th, [[1, 0, 0]], [1.0], (0.0, 120.0), kind, 3, 3, drop, ts,
                            fsptol)
    assert len(out) == len(ts) + 1
    assert out.t[-1] == out.t[-2] == 120.0 and np.array_equal(out.p[-1], out.p[-2])
    assert len(out.events) >= 2
    for t, p, s in zip(out.t, out.p, out.sinks):
        assert abs(p.sum() + s.sum() - 1.0) < 1.5e-8
        assert s.sum() <= fsptol * t / 120.0 + 1e-12


def test_solve_adaptive_matches_big_fixed_box():
    # the adaptive solution restricted to its space equals the fixed solve on a box that
    # holds it, up to the truncation the budget allows
    S, props, th = telegraph((0.05, 0.1, 5.0, 1.0))
    fsptol = 1e-6
    ts = [0.0, 10.0, 30.0]
    out = OA.solve_adaptive(S, to_oracle(props), th, [[1, 0, 0]], [1.0], (0.0, 30.0), OA.RSTEP, 5, 5, False, ts,
                            fsptol)
    box = [[1, 0, m] for m in range(61)] + [[0, 1, m] for m in range(61)]
    x0 = np.zeros(len(box))
    x0[0] = 1.0
    ref = solve_fixed(S, to_oracle(props), th, box, x0, ts)
    index = {tuple(x): i for i, x in enumerate(box)}
    for k in range(len(ts)):
        full = ref.p[k]
        sub = np.array([full[index[tuple(x)]] for x in out.states[k]])
        assert np.abs(sub - out.p[k]).sum() <= fsptol * ts[k] / 30.0 + 1e-9


def test_solve_adaptive_time_varying():
    S, props, th = telegraph_tv("sep")
    out = OA.solve_adaptive(S, to_oracle(props), th, [[1, 0, 0]], [1.0], (0.0, 20.0), OA.SELECTIVE, 10, 10, True,
                            [0.0, 10.0, 20.0], 1e-4)
    assert len(out) == 4
    for t, p, s in zip(out.t, out.p, out.sinks):
        assert abs(p.sum() + s.sum() - 1.0) < 1.5e-8
