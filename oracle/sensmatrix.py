"""Oracle restatement of NumCME.jl's forward-sensitivity FSP matrix and fixed-space
sensitivity solve (TEST INFRASTRUCTURE ONLY).

Restates

* ``ForwardSensFspMatrixSparse{RealT}(model, statespace)``
  (``src/forwardsensfspmatrix/forwardsensfspmatrixsparse/sensfspmatrixsparse.jl:31-95``):
  one time-invariant derivative matrix per parameter summed over the reactions that
  depend on it (``:43-58``); per (separable reaction, parameter) pair the matrix of the
  state factor's derivative plus a link to the reaction's factor matrix (``:60-71``); per
  (joint reaction, parameter) pair a pattern refreshed at every ``t`` (``:73-82``);
* its ``matvec!`` (``:97-142``): block 0 is ``A v0``; block p is
  ``A v_p + dA_p v0 + sum_sep [tf_r(t) dSF_rp + dtf_rp(t) F_r] v0 + sum_joint J_rp(t) v0``;
* the fixed-space form of the sensitivity ``solve``
  (``src/forwardsenscme/sparse/forwardsenscmesparse.jl:99-215`` with no adaptation):
  ODE state ``[p; sinks; S_1; dsinks_1; ...; S_P; dsinks_P]`` (block stride n,
  ``:132-138``), outputs ``t, p, sinks, S, dsinks`` per saved time (``:168-174``).

Propensity partial derivatives are given by the caller as plain callables (the reference
obtains them with ForwardDiff, ``src/cmemodel/senstools/forwarddiffpropensity.jl``):
``grads[(r, p)]`` is ``DTimeInvariant(f)`` (``f(x, theta)``), ``DSeparable(dtf, dsf)``
(``dtf(t, theta)`` and ``dsf(x, theta)``, either may be None for zero) or ``DJoint(f)``
(``f(t, x, theta)``); the keys are the gradient sparsity pattern. The integrator is
SciPy's DOP853 at rtol 1e-12.

Parity pins (``tests/test_oracle_sens.py``): the reference's analytic telegraph
derivative matrices (``test/sensmat/telegraph.jl:87-188``), zero column sums
(``telegraph.jl:33-43``, ``poisson.jl:6-21``), the hand derivatives against central
differences of the propensities, ``dA_p v`` against central differences of the oracle
mat-vec in theta, and the sensitivity solve against central differences of the
probability solve.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
from scipy.integrate import solve_ivp

from .fspmatrix import FspMatrix, _cols, _csc, _eval_state, generate_entries


class DTimeInvariant:
    kind = "ti"

    def __init__(self, f):
        self.f = f


class DSeparable:
    kind = "sep"

    def __init__(self, dtfactor=None, dstatefactor=None):
        self.dtfactor = dtfactor
        self.dstatefactor = dstatefactor


class DJoint:
    kind = "joint"

    def __init__(self, f):
        self.f = f


class SensFspMatrix:
    """Restatement of ``ForwardSensFspMatrixSparse`` (sensfspmatrixsparse.jl:9-95)."""

    def __init__(self, space, propensities, parameters, grads):
        self.A = FspMatrix(space, propensities, parameters)
        self.theta = np.asarray(parameters, dtype=np.float64)
        self.P = len(self.theta)
        self.props = list(propensities)
        self.states = space.states.copy()
        n = self.A.rowcount
        self.n = n
        # time-invariant reactions: one summed matrix per parameter (:43-58)
        self.ti_dA = []
        for ip in range(self.P):
            R, C, V = [], [], []
            for r in self.A.ti_ids:
                g = grads.get((r, ip))
                if g is None:
                    continue
                rr, cc, vv = generate_entries(space, _eval_state(g.f, self.states, self.theta), r)
                R.append(rr), C.append(cc), V.append(vv)
            if R:
                self.ti_dA.append(_csc(np.concatenate(R), np.concatenate(C), np.concatenate(V), n))
            else:
                self.ti_dA.append(sp.csc_matrix((n, n)))
        # separable reactions (:60-71): (ip, r, dSF matrix, position of r's factor matrix)
        self.sep = []
        for ip in range(self.P):
            for j, r in enumerate(self.A.sep_ids):
                g = grads.get((r, ip))
                if g is None:
                    continue
                dsf = np.zeros(space.count) if g.dstatefactor is None else _eval_state(g.dstatefactor, self.states,
                                                                                       self.theta)
                rr, cc, vv = generate_entries(space, dsf, r)
                self.sep.append((ip, r, _csc(rr, cc, vv, n), j, g.dtfactor))
        # joint reactions (:73-82): pattern, refreshed with the derivative at each t
        self.joint = []
        for ip in range(self.P):
            for r in self.A.joint_ids:
                g = grads.get((r, ip))
                if g is None:
                    continue
                rr, cc, vv = generate_entries(space, None, r)
                self.joint.append((ip, r, _csc(rr, cc, vv, n), g.f))

    def _refresh(self, M, f, t):
        # _update_sparsematrix! with the derivative (fspsparsematrix.jl:154-166)
        nstates = self.states.shape[0]
        val = np.broadcast_to(np.asarray(f(t, _cols(self.states), self.theta), dtype=np.float64), (nstates,))
        colid = np.repeat(np.arange(M.shape[1]), np.diff(M.indptr))
        v = np.where(colid < nstates, val[np.minimum(colid, nstates - 1)], 0.0)
        M.data[:] = np.where(M.indices == colid, -1.0 * v, v)

    def dA_matvec(self, t, ip, v0):
        """``dA_p(t) v0``: the parameter-p derivative terms of block p."""
        out = self.ti_dA[ip] @ v0
        for (jp, r, dsf_m, j, dtf) in self.sep:
            if jp != ip:
                continue
            out = out + self.props[r].tfactor(t, self.theta) * (dsf_m @ v0)
            if dtf is not None:
                out = out + float(dtf(t, self.theta)) * (self.A.sep[j] @ v0)
        for (jp, r, M, f) in self.joint:
            if jp != ip:
                continue
            self._refresh(M, f, t)
            out = out + M @ v0
        return out

    def matvec(self, t, vs):
        """``matvec!(out, t, SA, vs)`` (sensfspmatrixsparse.jl:97-142); block stride n."""
        n = self.n
        vs = np.asarray(vs, dtype=np.float64)
        out = np.zeros_like(vs)
        v0 = vs[:n]
        out[:n] = self.A.matvec(t, v0)
        for ip in range(self.P):
            blk = slice((ip + 1) * n, (ip + 2) * n)
            out[blk] = self.A.matvec(t, vs[blk]) + self.dA_matvec(t, ip, v0)
        return out


class SensOutput:
    """``ForwardSensFspOutputSparse`` (sensoutputsparse.jl:20-26) on a fixed space."""

    def __init__(self, states):
        self.states = states
        self.t, self.p, self.sinks, self.S, self.dsinks = [], [], [], [], []

    def __len__(self):
        return len(self.t)


def solve_sens_fixed(space, propensities, parameters, grads, p0, S0, saveat, rtol=1e-12, atol=1e-15,
                     breakpoints=()):
    """Sensitivity solve on the fixed space (forwardsenscmesparse.jl:99-215, no adaptation)."""
    SA = SensFspMatrix(space, propensities, parameters, grads)
    n, P, S = SA.n, SA.P, space.count
    u0 = np.zeros(n * (P + 1))
    u0[:S] = p0
    for ip in range(P):
        u0[(ip + 1) * n:(ip + 1) * n + S] = S0[ip]
    saveat = np.asarray(saveat, dtype=np.float64)
    pts = sorted(set([float(saveat[0]), float(saveat[-1])] +
                     [b for b in breakpoints if saveat[0] < b < saveat[-1]]))
    us = {0: u0.copy()} if saveat[0] == pts[0] else {}
    u = u0.copy()
    for a, b in zip(pts[:-1], pts[1:]):
        te = [float(t) for t in saveat if a < t <= b]
        ev = sorted(set(te + [b]))
        sol = solve_ivp(lambda t, y: SA.matvec(t, y), (a, b), u, method="DOP853", rtol=rtol, atol=atol, t_eval=ev)
        if not sol.success:
            raise RuntimeError(sol.message)
        got = {float(t): k for k, t in enumerate(sol.t)}
        for t in te:
            us[int(np.nonzero(saveat == t)[0][0])] = sol.y[:, got[t]].copy()
        u = sol.y[:, got[b]].copy()
    out = SensOutput(space.states.copy())
    for k, t in enumerate(saveat):
        y = us[k]
        out.t.append(float(t))
        out.p.append(y[:S].copy())
        out.sinks.append(y[S:n].copy())
        out.S.append([y[(ip + 1) * n:(ip + 1) * n + S].copy() for ip in range(P)])
        out.dsinks.append([y[(ip + 1) * n + S:(ip + 2) * n].copy() for ip in range(P)])
    return out
