"""Pins of the oracle's fixed-space solve (CPU only).

* closed-form telegraph moments: P(G_on), E[m], E[m G_on], E[m^2] obey a linear ODE
  with constant forcing, solved here by an augmented 5x5 matrix exponential; the FSP
  solution on 0..399 mRNA matches them to truncation level;
* mass conservation sum(p) + sum(sinks) = 1 (test/test_solver.jl:63-65) and one snapshot
  per saved time (:77);
* the dense-expm, expm_multiply and DOP853 paths agree;
* the FSP bound: a smaller box solution embedded in a larger one differs by at most the
  smaller box's sink mass (Munsky & Khammash 2006).
"""
import numpy as np
import scipy.linalg as sla

import models
from oracle.solve import solve_fixed


def telegraph_moments(th, ts):
    k_on, k_off, k_r, gam = th
    # y = [g, mu, nu, m2, 1]; d/dt y = M y
    M = np.zeros((5, 5))
    M[0, 0], M[0, 4] = -(k_on + k_off), k_on
    M[1, 0], M[1, 1] = k_r, -gam
    M[2, 0], M[2, 1], M[2, 2] = k_r, k_on, -(k_on + k_off + gam)
    M[3, 0], M[3, 1], M[3, 2], M[3, 3] = k_r, gam, 2 * k_r, -2 * gam
    y0 = np.array([0.0, 0.0, 0.0, 0.0, 1.0])
    return np.array([sla.expm(M * t) @ y0 for t in ts])


def test_telegraph_moments_c1():
    S, props, th = models.telegraph()
    states = models.telegraph_states(399)
    p0 = np.zeros(800)
    p0[0] = 1.0
    ts = np.arange(0, 201, 10.0)
    out = solve_fixed(S, models.to_oracle(props), th, states, p0, ts)
    X = np.asarray(states)
    ref = telegraph_moments(th, ts)
    assert len(out) == len(ts)
    for k, t in enumerate(ts):
        p = out.p[k]
        g = p[X[:, 1] == 1].sum()
        mu = (p * X[:, 2]).sum()
        nu = (p * X[:, 2] * X[:, 1]).sum()
        m2 = (p * X[:, 2] ** 2).sum()
        assert abs(g - ref[k, 0]) < 1e-12
        assert abs(mu - ref[k, 1]) < 1e-9 * max(1, ref[k, 1])
        assert abs(nu - ref[k, 2]) < 1e-9 * max(1, ref[k, 2])
        assert abs(m2 - ref[k, 3]) < 1e-9 * max(1, ref[k, 3])
        assert abs(p.sum() + out.sinks[k].sum() - 1.0) < 1e-12


def test_paths_agree_toggle_small():
    S, props, th = models.toggle()
    states = models.box_states([30, 30])
    p0 = np.zeros(len(states))
    p0[0] = 1.0
    ts = np.array([0.0, 0.5, 1.0, 2.0])
    a = solve_fixed(S, models.to_oracle(props), th, states, p0, ts, method="expm_multiply")
    b = solve_fixed(S, models.to_oracle(props), th, states, p0, ts, method="dense")
    c = solve_fixed(S, models.to_oracle(props), th, states, p0, ts, method="ode", rtol=1e-12, atol=1e-16)
    for k in range(len(ts)):
        assert np.abs(a.p[k] - b.p[k]).sum() < 1e-11
        assert np.abs(c.p[k] - b.p[k]).sum() < 1e-9
        assert abs(a.p[k].sum() + a.sinks[k].sum() - 1) < 1e-12


def test_fsp_bound_nested_boxes():
    S, props, th = models.toggle()
    small, big = models.box_states([40, 40]), models.box_states([60, 60])
    ts = np.array([0.0, 1.0, 3.0])
    ps, pb = np.zeros(len(small)), np.zeros(len(big))
    ps[0] = pb[0] = 1.0
    a = solve_fixed(S, models.to_oracle(props), th, small, ps, ts)
    b = solve_fixed(S, models.to_oracle(props), th, big, pb, ts)
    idx = {tuple(x): i for i, x in enumerate(big.tolist())}
    emb = np.array([idx[tuple(x)] for x in small.tolist()])
    for k in range(len(ts)):
        diff = np.abs(b.p[k][emb] - a.p[k]).sum() + np.delete(b.p[k], emb).sum()
        assert np.all(b.p[k][emb] >= a.p[k] - 1e-14)  # FSP lower bound
        assert diff <= a.sinks[k].sum() + 1e-12


def test_time_varying_mass_conservation():
    S, props, th = models.telegraph_sens()
    states = models.telegraph_states(60)
    p0 = np.zeros(len(states))
    p0[0] = 1.0
    ts = np.array([0.0, 5.0, 10.0, 30.0])
    kinks = [20.0 * (0.5 + 2 * k) for k in range(3)]  # sin(pi t/L) = 1 where 1-sin touches 0
    out = solve_fixed(S, models.to_oracle(props), th, states, p0, ts, breakpoints=kinks)
    for k in range(len(ts)):
        assert abs(out.p[k].sum() + out.sinks[k].sum() - 1) < 1e-10


def _toggle_generator(bounds, th=None):
    from oracle.fspmatrix import FspMatrix
    from oracle.statespace import StateSpace
    S, props, th0 = models.toggle() if th is None else models.toggle(th)
    A = FspMatrix(StateSpace(S, models.box_states(bounds)), models.to_oracle(props), th0).to_csc(0.0)
    u0 = np.zeros(A.shape[0])
    u0[0] = 1.0
    return A, u0


def test_uniformization_vs_dense_expm_and_expm_multiply():
    # the exact method of the large batched parity cases, pinned by two independent exact ones
    from oracle import cpu_ref
    from oracle.solve import uniformization
    import scipy.sparse.linalg as spla

    ts = np.arange(0.0, 6.0)
    A, u0 = _toggle_generator([24, 24])
    U = uniformization(A, u0, ts)
    Uc = cpu_ref.uniformization(A, u0, ts)
    for k, t in enumerate(ts):
        ref = sla.expm(A.toarray() * t) @ u0
        assert np.abs(U[k] - ref).sum() < 1e-12
        assert np.abs(Uc[k] - U[k]).sum() < 1e-14
    # a slow, widely spread instance (gamma = 0.05) against expm_multiply
    A, u0 = _toggle_generator([119, 119], (30.0, 10.0, 2.5, 1.0, 0.05))
    Uc = cpu_ref.uniformization(A, u0, ts)
    u = u0
    for k in range(1, len(ts)):
        u = spla.expm_multiply(A * (ts[k] - ts[k - 1]), u)
        assert np.abs(Uc[k] - u).sum() < 1e-11
        assert abs(Uc[k].sum() - 1.0) < 1e-12  # the sinks hold the rest (none reached here)


def test_c_krylov_restatement_matches_exact():
    # the timed CPU restatement (one-thread CSC scatter, all-thread CSR) computes the solution
    from oracle import cpu_ref

    ts = np.array([0.0, 0.5, 2.0])
    A, u0 = _toggle_generator([40, 40])
    ref = cpu_ref.uniformization(A, u0, ts)
    for layout, th in (("csc", 1), ("csr", cpu_ref.cpu_count())):
        W, steps, _ = cpu_ref.Expv(A, m=30, tol=1e-12, layout=layout, threads=th).solve(u0, ts)
        assert steps > 0
        for k in range(len(ts)):
            assert np.abs(W[k] - ref[k]).sum() < 1e-10
