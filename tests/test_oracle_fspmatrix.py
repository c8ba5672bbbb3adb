"""Pins of the oracle's FSP matrix restatement (CPU only).

* exact nonzero counts of the two fixed configurations (explicit zeros included):
  telegraph on 800 states -> 3,198 and toggle on 1024^2 states -> 5,240,832;
* element-wise equality with the analytic telegraph generator built by the rules of
  test/sensmat/telegraph.jl:46-85 (an independent, state-formula construction);
* a brute-force dense generator built from the CME definition on small spaces;
* zero column sums and separable == joint at t = 0 (test/test_fspmat.jl:39-68).
"""
import numpy as np
import pytest

import models
from oracle.fspmatrix import FspMatrix
from oracle.statespace import StateSpace

EPS = np.finfo(float).eps


def oracle_matrix(S, props, theta, states):
    sp = StateSpace(S, states)
    return sp, FspMatrix(sp, models.to_oracle(props), theta)


def test_nnz_telegraph_800():
    S, props, th = models.telegraph()
    _, A = oracle_matrix(S, props, th, models.telegraph_states(399))
    assert A.shape == (804, 804)
    assert A.nnz() == 3198


def test_nnz_toggle_1024():
    S, props, th = models.toggle()
    _, A = oracle_matrix(S, props, th, models.box_states([1023, 1023]))
    assert A.shape == (1048580, 1048580)
    assert A.nnz() == 5240832


def analytic_telegraph(nmax, th):
    """Generator by the construction rules of test/sensmat/telegraph.jl:46-85 (0-based)."""
    k01, k10, lam, gam = th
    n = 2 * (nmax + 1) + 4
    A = np.zeros((n, n))
    for gene in (0, 1):
        for m in range(nmax + 1):
            x = (1 - gene, gene, m)
            a1, a2, a3, a4 = k01 * x[0], k10 * x[1], lam * x[1], gam * x[2]
            idx = gene * (nmax + 1) + m
            A[idx, idx] += -(a1 + a2 + a3 + a4)
            if gene == 0:
                A[nmax + 1 + m, idx] += a1
            if gene == 1:
                A[m, idx] += a2
                A[(nmax + 1 + m + 1) if m < nmax else 2 * (nmax + 1) + 2, idx] += a3
            if m > 0:
                A[gene * (nmax + 1) + m - 1, idx] += a4
    return A


@pytest.mark.parametrize("nmax", [5, 60, 399])
def test_analytic_telegraph(nmax):
    S, props, th = models.telegraph()
    _, A = oracle_matrix(S, props, th, models.telegraph_states(nmax))
    ref = analytic_telegraph(nmax, th)
    got = A.to_csc(0.0).toarray()
    assert np.abs(got - ref).max() <= 8 * EPS * np.abs(ref).max()


def brute_generator(S, props, th, states, t=0.0):
    """Q[i, j] = rate j -> i from the CME definition; sink row S+r collects exits via r."""
    states = [tuple(s) for s in states]
    index = {x: i for i, x in enumerate(states)}
    n, R = len(states), S.shape[1]
    Q = np.zeros((n + R, n + R))
    for j, x in enumerate(states):
        cols = [np.array([v]) for v in x]
        for r, pr in enumerate(props):
            if pr[0] == "ti":
                a = float(np.asarray(pr[1](cols, th)).ravel()[0])
            elif pr[0] == "sep":
                a = float(pr[1](t, th)) * float(np.asarray(pr[2](cols, th)).ravel()[0])
            else:
                a = float(np.asarray(pr[1](t, cols, th)).ravel()[0])
            Q[j, j] -= a
            y = tuple(int(v) for v in np.asarray(x) + S[:, r])
            if y in index:
                Q[index[y], j] += a
            elif min(y) >= 0:
                Q[n + r, j] += a
    return Q


@pytest.mark.parametrize("which", ["toggle", "telegraph", "gene4"])
def test_brute_force_generator(which):
    if which == "toggle":
        S, props, th = models.toggle()
        sp = StateSpace(S, [[0, 0]])
        sp.expand(6)
    elif which == "telegraph":
        S, props, th = models.telegraph()
        sp = StateSpace(S, [[1, 0, 0]])
        sp.expand(7)
    else:
        S, props, th = models.gene4()
        sp = StateSpace(S, [[1, 0, 0, 0, 0, 0]])
        sp.expand(4)
    A = FspMatrix(sp, models.to_oracle(props), th)
    for t in (0.0, 37.5):
        ref = brute_generator(S, props, th, sp.states.tolist(), t)
        got = A.to_csc(t).toarray()
        assert np.abs(got - ref).max() <= 16 * EPS * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("kind", ["ti", "sep", "joint"])
def test_zero_column_sums(kind):
    # test_fspmat.jl:126-154: sum(A*ones) == 0 at t = 0 and t = 1
    if kind == "ti":
        S, props, th = models.telegraph((0.05, 0.1, 5.0, 1.0))
    else:
        S, props, th = models.telegraph_tv(kind)
    sp = StateSpace(S, [[1, 0, 0]])
    sp.expand(2)
    A = FspMatrix(sp, models.to_oracle(props), th)
    for t in (1.0, 0.0):
        w = A.matvec(t, np.ones(A.rowcount))
        assert abs(w.sum()) <= 1e-14


def test_separable_equals_joint_at_t0():
    # test_fspmat.jl:155: norm(w1 - w2) == 0 with the time factor equal to 1 at t = 0
    S, sep, th = models.telegraph_tv("sep")
    _, joint, _ = models.telegraph_tv("joint")
    sp = StateSpace(S, [[1, 0, 0]])
    sp.expand(2)
    v = np.ones(sp.count + 4)
    w1 = FspMatrix(sp, models.to_oracle(sep), th).matvec(0.0, v)
    w2 = FspMatrix(sp, models.to_oracle(joint), th).matvec(0.0, v)
    assert np.array_equal(w1, w2)


def test_joint_refresh_tracks_time():
    # _update_sparsematrix! runs only when t changes (fspsparsematrix.jl:206-213)
    S, joint, th = models.telegraph_tv("joint")
    _, sep, _ = models.telegraph_tv("sep")
    sp = StateSpace(S, [[1, 0, 0]])
    sp.expand(3)
    Aj = FspMatrix(sp, models.to_oracle(joint), th)
    As = FspMatrix(sp, models.to_oracle(sep), th)
    v = np.linspace(0.1, 1.0, sp.count + 4)
    for t in (0.3, 0.3, 1.7, 0.3):
        assert np.allclose(Aj.matvec(t, v), As.matvec(t, v), rtol=1e-13, atol=1e-15)
