"""Oracle restatement of NumCME.jl's FSP ``solve`` (TEST INFRASTRUCTURE ONLY).

Restates the fixed-space solve of ``src/transientcme/sparse/fspsolve.jl:10-41``:
the state space is the initial distribution's state list in order (``:16``),
the matrix is FP64 (``:20``), the ODE state is ``[p0; zeros(R)]`` (``:21``) and one
snapshot ``(t, p, sinks)`` is produced per saved time (``:29-38``).

The reference integrates with CVODE (third-party, unpinned). The oracle
integrates *exactly* instead, because the FSP solution is unique:

* time-invariant generators: ``scipy.sparse.linalg.expm_multiply``
  (Al-Mohy & Higham 2011) or, for small spaces, dense ``scipy.linalg.expm``;
* time-varying generators: ``scipy.integrate.solve_ivp`` DOP853 at
  ``rtol=1e-12, atol=1e-15`` with the ``max(0, .)`` kink times as break points.

Parity pins (``tests/test_oracle_solve.py``): mass conservation
``sum(p)+sum(sinks) = 1`` (``test/test_solver.jl:63-65``), snapshot count ==
``len(saveat)`` (``:77``), the closed-form telegraph moment ODEs (mean mRNA and
gene-state probabilities), and agreement of the dense-``expm`` and
``expm_multiply`` paths.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla
import scipy.sparse.linalg as spla
from scipy.integrate import solve_ivp

from .fspmatrix import FspMatrix
from .statespace import StateSpace


class Output:
    """``FspOutputSparse`` (fspoutput.jl:19-74): lists ``t``, ``p`` (value arrays), ``sinks``."""

    def __init__(self, states):
        self.states = states
        self.t, self.p, self.sinks = [], [], []

    def __len__(self):
        return len(self.t)


def solve_fixed(stoich, propensities, parameters, init_states, init_values, saveat,
                method="auto", rtol=1e-12, atol=1e-15, breakpoints=()):
    """Fixed-space FSP solve (fspsolve.jl:10-41) with an exact integrator."""
    space = StateSpace(stoich, init_states)
    A = FspMatrix(space, propensities, parameters)
    S = space.count
    R = space.sink_count
    u0 = np.concatenate([np.asarray(init_values, dtype=np.float64), np.zeros(R)])
    saveat = np.asarray(saveat, dtype=np.float64)
    out = Output(space.states.copy())
    tv = bool(A.sep_ids or A.joint_ids)
    if method == "auto":
        method = "ode" if tv else ("dense" if A.rowcount <= 2000 else "expm_multiply")
    if method == "dense":
        M = A.to_csc(0.0).toarray()
        us = [sla.expm(M * (t - saveat[0])) @ u0 for t in saveat]
    elif method == "expm_multiply":
        M = A.to_csc(0.0)
        us, u, tprev = [], u0, saveat[0]
        for t in saveat:
            if t > tprev:
                u = spla.expm_multiply(M * (t - tprev), u)
            us.append(u.copy())
            tprev = t
    else:
        us = _integrate_ode(A, u0, saveat, rtol, atol, breakpoints)
    for t, u in zip(saveat, us):
        out.t.append(float(t))
        out.p.append(u[:S].copy())
        out.sinks.append(u[S:].copy())
    return out


def _integrate_ode(A, u0, saveat, rtol, atol, breakpoints):
    pts = sorted(set([float(saveat[0]), float(saveat[-1])] + [b for b in breakpoints if saveat[0] < b < saveat[-1]]))
    us = {}
    u = u0.copy()
    if saveat[0] == pts[0]:
        us[0] = u0.copy()
    for a, b in zip(pts[:-1], pts[1:]):
        te = [t for t in saveat if a < t <= b]
        sol = solve_ivp(lambda t, y: A.matvec(t, y), (a, b), u, method="DOP853", rtol=rtol, atol=atol,
                        t_eval=te if te else None, dense_output=False)
        if not sol.success:
            raise RuntimeError(sol.message)
        for k, t in enumerate(te):
            us[int(np.nonzero(saveat == t)[0][0])] = sol.y[:, k].copy()
        u = sol.y[:, -1].copy() if te and te[-1] == b else _endpoint(A, u, a, b, rtol, atol)
    return [us[i] for i in range(len(saveat))]


def _endpoint(A, u, a, b, rtol, atol):
    sol = solve_ivp(lambda t, y: A.matvec(t, y), (a, b), u, method="DOP853", rtol=rtol, atol=atol)
    return sol.y[:, -1].copy()


def expm_dense(M):
    """Dense matrix exponential (scipy, Pade 13 scaling-and-squaring; Al-Mohy & Higham 2009)."""
    return sla.expm(np.asarray(M, dtype=np.float64))


def uniformization(A, u0, ts, tail=1e-14):
    """Exact transient solution of du/dt = A u for a generator A (non-negative off-diagonal,
    non-positive diagonal, columns summing to <= 0) at the ascending times ``ts`` (from
    ``ts[0]``), by uniformization (Jensen 1953; Grassmann 1977): with L >= max |A_jj| and
    P = I + A / L (a sub-stochastic matrix), u(t) = sum_k e^{-L t} (L t)^k / k! P^k u0.
    Every term is non-negative for a non-negative u0, so there is no cancellation; the
    series is cut where the Poisson weights left out sum to less than ``tail``. Intervals
    between output times are advanced one after another.
    """
    import math

    import scipy.sparse as sps
    from scipy.stats import poisson

    A = sps.csr_matrix(A)
    L = float(np.abs(A.diagonal()).max()) * 1.0000001 or 1.0
    P = sps.identity(A.shape[0], format="csr") + A / L
    P = sps.csr_matrix(P)
    u = np.asarray(u0, dtype=np.float64).copy()
    out = [u.copy()]
    for a, b in zip(ts[:-1], ts[1:]):
        lt = L * (b - a)
        if lt == 0.0:
            out.append(u.copy())
            continue
        kmax = int(poisson.isf(tail, lt)) + 2
        # Poisson weights by the recurrence w_k = w_{k-1} lt / k, scaled from the mode down
        # and up so that no weight underflows before it is used
        mode = int(lt)
        logw_mode = -lt + mode * math.log(lt) - math.lgamma(mode + 1)
        w = np.empty(kmax + 1)
        w[mode] = 1.0 if mode <= kmax else 0.0
        for k in range(mode + 1, kmax + 1):
            w[k] = w[k - 1] * lt / k
        for k in range(min(mode, kmax + 1) - 1, -1, -1):
            w[k] = w[k + 1] * (k + 1) / lt
        w *= math.exp(logw_mode)
        acc = w[0] * u
        term = u
        for k in range(1, kmax + 1):
            term = P @ term
            if w[k] != 0.0:
                acc += w[k] * term
        u = acc
        out.append(u.copy())
    return out
