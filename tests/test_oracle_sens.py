"""Pins of the oracle's forward-sensitivity restatement (CPU only).

References: src/forwardsensfspmatrix/forwardsensfspmatrixsparse/sensfspmatrixsparse.jl,
test/sensmat/telegraph.jl (analytic derivative matrices, zero sums), test/sensmat/poisson.jl,
src/forwardsenscme/sparse/forwardsenscmesparse.jl (fixed-space solve).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import models
from oracle.fspmatrix import FspMatrix, TimeInvariant
from oracle.sensmatrix import DTimeInvariant, SensFspMatrix, solve_sens_fixed
from oracle.solve import solve_fixed
from oracle.statespace import StateSpace


def _prop_value(pr, t, states, theta):
    cols = [states[:, s].astype(float) for s in range(states.shape[1])]
    n = states.shape[0]
    if pr[0] == "ti":
        v = pr[1](cols, theta)
    elif pr[0] == "sep":
        v = pr[1](t, theta) * np.asarray(pr[2](cols, theta))
    else:
        v = pr[1](t, cols, theta)
    return np.broadcast_to(np.asarray(v, dtype=float), (n,))


def _grad_value(g, pr, t, states, theta):
    cols = [states[:, s].astype(float) for s in range(states.shape[1])]
    n = states.shape[0]
    if g.kind == "ti":
        v = g.f(cols, theta)
    elif g.kind == "sep":
        sf = np.asarray(pr[2](cols, theta), dtype=float)
        tf = pr[1](t, theta)
        v = 0.0
        if g.dtfactor is not None:
            v = v + g.dtfactor(t, theta) * sf
        if g.dstatefactor is not None:
            v = v + tf * np.asarray(g.dstatefactor(cols, theta), dtype=float)
    else:
        v = g.f(t, cols, theta)
    return np.broadcast_to(np.asarray(v, dtype=float), (n,))


CASES = {
    "telegraph_sens": (models.telegraph_sens, models.telegraph_sens_grads, models.telegraph_states(30)),
    "gene4": (models.gene4, models.gene4_grads, models.gene4_states(5)),
    "toggle": (models.toggle, models.toggle_grads, models.box_states([40, 40])),
    "twelve_slot": (models.twelve_slot, models.twelve_slot_grads, models.box_states([6, 5, 4])),
}


@pytest.mark.parametrize("case", list(CASES))
def test_hand_gradients_match_central_differences(case):
    mk, gk, states = CASES[case]
    S, props, th = mk()
    grads = gk()
    states = np.asarray(states)
    for t in (0.0, 7.5, 33.0, 150.0):
        for r, pr in enumerate(props):
            for p in range(len(th)):
                h = 1e-6 * max(1.0, abs(th[p]))
                tp, tm = th.copy(), th.copy()
                tp[p] += h
                tm[p] -= h
                fd = (_prop_value(pr, t, states, tp) - _prop_value(pr, t, states, tm)) / (2 * h)
                if (r, p) not in grads:
                    assert np.abs(fd).max() == 0.0, f"({r},{p}) missing from the pattern"
                    continue
                an = _grad_value(grads[(r, p)], pr, t, states, th)
                scale = max(1.0, np.abs(an).max())
                assert np.abs(an - fd).max() <= 1e-6 * scale, (case, r, p, t)


# ---- the reference's analytic telegraph matrices (test/sensmat/telegraph.jl:46-188) ----
def _analytic(nmax, t, th):
    k01, k10, lam, gam, L = th
    ns = 2 * (nmax + 1)
    n = ns + 4
    tf = max(0.0, 1.0 - np.sin(np.pi * t / L))
    on = 1.0 - np.sin(np.pi * t / L) >= 0
    dtf_dL = (np.cos(np.pi * t / L) * np.pi * t / L ** 2) if on else 0.0

    def M():
        return sp.lil_matrix((n, n))
    A, dk01, dk10, dlam, dgam, dL = M(), M(), M(), M(), M(), M()
    for gene in (0, 1):
        for m in range(nmax + 1):
            i = gene * (nmax + 1) + m
            a1 = k01 * (1 - gene)
            a2 = k10 * gene
            a3 = lam * gene
            a4 = tf * gam * m
            A[i, i] += -(a1 + a2 + a3 + a4)
            if gene == 0:
                A[nmax + 1 + m, i] += a1
                dk01[i, i] += -1.0
                dk01[nmax + 1 + m, i] += 1.0
            else:
                A[m, i] += a2
                dk10[i, i] += -1.0
                dk10[m, i] += 1.0
                succ = nmax + 1 + m + 1 if m < nmax else ns + 2
                A[succ, i] += a3
                dlam[i, i] += -1.0
                dlam[succ, i] += 1.0
            if m > 0:
                A[i - 1, i] += a4
                dgam[i, i] += -tf * m
                dgam[i - 1, i] += tf * m
                dL[i, i] += -gam * m * dtf_dL
                dL[i - 1, i] += gam * m * dtf_dL
    return A.tocsr(), [dk01.tocsr(), dk10.tocsr(), dlam.tocsr(), dgam.tocsr(), dL.tocsr()]


@pytest.mark.parametrize("t", [10.0, 20.0, 30.0, 100.0])
def test_matvec_matches_reference_analytic_telegraph(t):
    nmax = 500
    S, props, th = models.telegraph_sens()
    space = StateSpace(S, models.telegraph_states(nmax))
    SA = SensFspMatrix(space, models.to_oracle(props), th, models.telegraph_sens_grads())
    n = SA.n
    v = np.ones(n * 6)
    v /= v.sum()
    out = SA.matvec(t, v)
    A, dAs = _analytic(nmax, t, th)
    ref = np.zeros_like(v)
    ref[:n] = A @ v[:n]
    for ip in range(5):
        ref[(ip + 1) * n:(ip + 2) * n] = dAs[ip] @ v[:n] + A @ v[(ip + 1) * n:(ip + 2) * n]
    assert np.abs(out - ref).max() <= n * np.finfo(float).eps


def test_zero_sums():
    # telegraph.jl:33-43 and poisson.jl:6-21
    S, props, th = models.telegraph_sens()
    space = StateSpace(S, models.telegraph_states(500))
    SA = SensFspMatrix(space, models.to_oracle(props), th, models.telegraph_sens_grads())
    n = SA.n
    v = np.ones(n * 6)
    out = SA.matvec(20.0, v)
    assert abs(out.sum() / v.sum()) <= n * 6 * np.finfo(float).eps
    Sp = np.array([[1, -1]])
    space = StateSpace(Sp, [[i] for i in range(1, 101)])
    props = [TimeInvariant(lambda x, p: p[0] + 0.0 * x[0]), TimeInvariant(lambda x, p: p[1] + 0.0 * x[0])]
    grads = {(0, 0): DTimeInvariant(lambda x, p: 1.0 + 0.0 * x[0]), (1, 1): DTimeInvariant(lambda x, p: 1.0 + 0.0 * x[0])}
    SA = SensFspMatrix(space, props, [10.0, 5.0], grads)
    out = SA.matvec(0.0, np.ones(3 * SA.n))
    assert abs(out.sum()) <= 1e-9


@pytest.mark.parametrize("case", ["toggle", "gene4", "telegraph_sens"])
def test_dA_matches_central_difference_of_matvec(case):
    mk, gk, states = CASES[case]
    S, props, th = mk()
    space = StateSpace(S, states)
    oprops = models.to_oracle(props)
    SA = SensFspMatrix(space, oprops, th, gk())
    v = np.random.default_rng(3).random(SA.n)
    for t in (0.0, 20.0, 300.0):
        for ip in range(len(th)):
            h = 1e-6 * max(1.0, abs(th[ip]))
            tp, tm = th.copy(), th.copy()
            tp[ip] += h
            tm[ip] -= h
            fd = (FspMatrix(space, oprops, tp).matvec(t, v) - FspMatrix(space, oprops, tm).matvec(t, v)) / (2 * h)
            an = SA.dA_matvec(t, ip, v)
            assert np.abs(an - fd).max() <= 1e-6 * max(1.0, np.abs(an).max()), (case, ip, t)


def test_sens_solve_matches_central_difference_of_solve():
    S, props, th = models.telegraph_sens((0.05, 0.1, 5.0, 0.5, 20.0))
    states = models.telegraph_states(40)
    space = StateSpace(S, states)
    oprops = models.to_oracle(props)
    p0 = np.zeros(len(states))
    p0[0] = 1.0
    ts = np.array([0.0, 5.0, 12.0, 25.0])
    # kinks of max(0, 1 - sin(pi t / L)) where sin = 1: t = L/2 + 2kL
    sol = solve_sens_fixed(space, oprops, th, models.telegraph_sens_grads(), p0, [np.zeros(len(states))] * 5, ts,
                           breakpoints=[10.0])
    for ip in range(5):
        h = 1e-5 * max(1.0, abs(th[ip]))
        tp, tm = th.copy(), th.copy()
        tp[ip] += h
        tm[ip] -= h
        rp = solve_fixed(S, oprops, tp, states, p0, ts, breakpoints=[10.0 * tp[4] / 20.0])
        rm = solve_fixed(S, oprops, tm, states, p0, ts, breakpoints=[10.0 * tm[4] / 20.0])
        for k in range(len(ts)):
            fd = (rp.p[k] - rm.p[k]) / (2 * h)
            dsink = (rp.sinks[k] - rm.sinks[k]) / (2 * h)
            scale = max(1e-3, np.abs(fd).max())
            assert np.abs(sol.S[k][ip] - fd).max() <= 1e-6 * scale + 1e-9, (ip, k)
            assert np.abs(sol.dsinks[k][ip] - dsink).max() <= 1e-6 * max(1e-3, np.abs(dsink).max()) + 1e-9
    for k in range(len(ts)):
        total = sum(s.sum() + d.sum() for s, d in zip(sol.S[k], sol.dsinks[k]))
        assert abs(total) <= 1e-9  # d/dtheta of the total mass is zero
