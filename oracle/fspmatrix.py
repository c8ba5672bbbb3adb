"""Oracle restatement of NumCME.jl's FSP matrix (TEST INFRASTRUCTURE ONLY).

Restates ``src/fspmatrix/sparse/fspsparsematrix.jl``:

* ``FspMatrixSparse{RealT}`` constructor       ``fspsparsematrix.jl:47-108``
* ``_generate_sparsematrix_entries``           ``fspsparsematrix.jl:120-152``
* ``_update_sparsematrix!``                    ``fspsparsematrix.jl:154-166``
* ``matvec!`` / ``matvecadd!``                 ``fspsparsematrix.jl:196-247``

and the propensity kinds of ``src/cmemodel/propensity.jl:50-153``.

Propensity callables are *vectorised*: ``x`` is a sequence of per-species
arrays (``x[s]`` holds species ``s`` over all states, 0-based species), ``p``
the parameter vector (0-based), ``t`` a float. A callable may return a scalar
(broadcast over states).

Parity pins (``tests/test_oracle_fspmatrix.py``): exact nnz counts of the two
fixed configurations (telegraph 800 states -> 3,198; toggle 1024^2 -> 5,240,832),
element-by-element equality with the analytic telegraph generator built by the
rules of ``test/sensmat/telegraph.jl:46-85``, a brute-force dense generator on
small spaces, zero column sums (``test/test_fspmat.jl:39-68``) and
separable == joint at ``t`` where the time factor is 1.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


class TimeInvariant:
    """``StandardTimeInvariantPropensity`` (propensity.jl:50-52): ``f(x, p)``."""

    kind = "ti"

    def __init__(self, f):
        self.f = f


class Separable:
    """``SeparableTimeVaryingPropensity`` (propensity.jl:91-94): ``tfactor(t,p)*statefactor(x,p)``."""

    kind = "sep"

    def __init__(self, tfactor, statefactor):
        self.tfactor = tfactor
        self.statefactor = statefactor


class Joint:
    """``JointTimeVaryingPropensity`` (propensity.jl:122-124): ``f(t, x, p)``."""

    kind = "joint"

    def __init__(self, f):
        self.f = f


def _cols(states):
    return [states[:, s] for s in range(states.shape[1])]


def _eval_state(f, states, theta):
    n = states.shape[0]
    v = np.asarray(f(_cols(states), theta), dtype=np.float64)
    return np.broadcast_to(v, (n,)).astype(np.float64)


def generate_entries(space, statefactor_values, r):
    """``_generate_sparsematrix_entries`` (fspsparsematrix.jl:120-152) for reaction ``r`` (0-based).

    ``statefactor_values`` is the per-state value array, or ``None`` for a joint
    propensity (pattern only, zero values). Returns 1-based (rows, cols, vals).
    """
    S = space.count
    rows = np.zeros(2 * S, dtype=np.int64)
    cols = np.zeros(2 * S, dtype=np.int64)
    vals = np.zeros(2 * S, dtype=np.float64)
    dval = np.zeros(S) if statefactor_values is None else statefactor_values
    idx = np.arange(1, S + 1)
    rows[:S] = idx
    cols[:S] = idx
    vals[:S] = -1.0 * dval
    sink = space.sink_conn[:, r]
    m = sink != 0
    cols[S + idx[m] - 1] = idx[m]
    rows[S + idx[m] - 1] = S + sink[m]
    vals[S + idx[m] - 1] = dval[m]
    cidx = space.state_conn[:, r]
    m2 = cidx != 0
    slot = cidx[m2] + S - 1
    rows[slot] = idx[m2]
    cols[slot] = cidx[m2]
    vals[slot] = -1.0 * vals[cidx[m2] - 1]
    keep = cols != 0
    return rows[keep], cols[keep], vals[keep]


def _csc(rows, cols, vals, n):
    # SparseArrays.sparse(I,J,V,m,n): duplicates summed, explicit zeros kept
    A = sp.coo_matrix((vals, (rows - 1, cols - 1)), shape=(n, n)).tocsc()
    A.sum_duplicates()
    return A


class FspMatrix:
    """Restatement of ``FspMatrixSparse`` (fspsparsematrix.jl:9-27,47-108)."""

    def __init__(self, space, propensities, parameters=()):
        self.parameters = np.asarray(parameters, dtype=np.float64)
        self.states = space.states.copy()
        S = space.count
        self.rowcount = S + space.sink_count
        self.propensities = list(propensities)
        self.ti_ids, self.sep_ids, self.joint_ids = [], [], []
        for r, a in enumerate(self.propensities):
            {"ti": self.ti_ids, "sep": self.sep_ids, "joint": self.joint_ids}[a.kind].append(r)
        n = self.rowcount
        if self.ti_ids:
            R, C, V = [], [], []
            for r in self.ti_ids:
                rr, cc, vv = generate_entries(space, _eval_state(self.propensities[r].f, self.states, self.parameters), r)
                R.append(rr), C.append(cc), V.append(vv)
            self.ti = _csc(np.concatenate(R), np.concatenate(C), np.concatenate(V), n)
        else:
            self.ti = None
        self.sep = []
        for r in self.sep_ids:
            rr, cc, vv = generate_entries(space, _eval_state(self.propensities[r].statefactor, self.states, self.parameters), r)
            self.sep.append(_csc(rr, cc, vv, n))
        self.joint = []
        for r in self.joint_ids:
            rr, cc, vv = generate_entries(space, None, r)
            self.joint.append(_csc(rr, cc, vv, n))
        self.t_cache = -np.inf

    @property
    def shape(self):
        return (self.rowcount, self.rowcount)

    def nnz(self):
        return (0 if self.ti is None else self.ti.nnz) + sum(m.nnz for m in self.sep) + sum(m.nnz for m in self.joint)

    def _update_joint(self, t):
        # _update_sparsematrix! (fspsparsematrix.jl:154-166)
        nstates = self.states.shape[0]
        for M, r in zip(self.joint, self.joint_ids):
            f = self.propensities[r].f
            val = np.broadcast_to(np.asarray(f(t, _cols(self.states), self.parameters), dtype=np.float64), (nstates,))
            colid = np.repeat(np.arange(M.shape[1]), np.diff(M.indptr))
            v = np.where(colid < nstates, val[np.minimum(colid, nstates - 1)], 0.0)
            M.data[:] = np.where(M.indices == colid, -1.0 * v, v)

    def matvecadd(self, out, t, v):
        """``matvecadd!`` (fspsparsematrix.jl:226-247): ``out += A(t) v``."""
        if self.ti is not None:
            out += self.ti @ v
        for M, r in zip(self.sep, self.sep_ids):
            out += self.propensities[r].tfactor(t, self.parameters) * (M @ v)
        if t != self.t_cache:
            self.t_cache = t
            self._update_joint(t)
        for M in self.joint:
            out += M @ v
        return out

    def matvec(self, t, v):
        """``matvec`` (fspsparsematrix.jl:196-217,254-258)."""
        out = np.zeros(self.rowcount)
        return self.matvecadd(out, t, np.asarray(v, dtype=np.float64))

    def to_csc(self, t):
        """The full generator ``A(t)`` as one CSC matrix (sum of all parts at time ``t``)."""
        A = sp.csc_matrix(self.shape)
        if self.ti is not None:
            A = A + self.ti
        for M, r in zip(self.sep, self.sep_ids):
            A = A + self.propensities[r].tfactor(t, self.parameters) * M
        if t != self.t_cache:
            self.t_cache = t
            self._update_joint(t)
        for M in self.joint:
            A = A + M
        return A.tocsc()
