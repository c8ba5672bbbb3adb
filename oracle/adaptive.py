"""Oracle restatement of NumCME.jl's adaptive FSP (TEST INFRASTRUCTURE ONLY).

Restates

* the drop selection of ``RStepAdapter`` / ``SelectiveRStepAdapter``
  (``src/transientcme/sparse/rstepadapters.jl:41-43`` with ``>=``, ``:94-96`` with ``>``):
  stable ``sortperm(p)``, ``dropcount = #{k : sum(p) - cumsum(p[perm])[k] >= thr}``,
  ``thr = 1 - t*fsptol/tend``, ``dropids = sort(perm[1:dropcount])``;
* ``init!`` / ``adapt!`` (``rstepadapters.jl:23-52, 74-110``), with the selective
  adapter's reaction list ``findall(dsinks .> 0)`` (``:104-105``);
* the adapt-and-restart driver (``fspsolve.jl:105-197``): sink-budget event
  ``sum(sinks) - fsptol*t/tend`` crossing zero from below, snapshot list with the
  saveat semantics of DifferentialEquations (start saved when it is a saveat point,
  then every saveat point up to the event or the end), the eps trick of ``:178-181`` and
  the duplicated final snapshot of ``:163-167`` / ``:189-194``.

Rounding conventions shared with the library (so that drop sets can be compared bit for
bit given the same ``p``): probabilities are ordered by Julia's ``isless`` (``-0.0``
before ``+0.0``), the cumulative sum uses a fixed association order (1024-element tiles
scanned Hillis-Steele style, tile totals scanned recursively, the prefix of the preceding
tiles added last), and ``sum(p)`` is taken as the last cumulative sum. The reference's
pairwise ``sum``/``cumsum`` differ from these by rounding only.

The integrator is SciPy's DOP853 at ``rtol=1e-12, atol=1e-15`` with its event root
finder; the reference uses DifferentialEquations callbacks (third-party, unpinned).

Parity pins (``tests/test_oracle_adaptive.py``): hand-derived drop sets that separate
``>=`` from ``>``; the literal ``sum(p) - cumsum(p[sortperm(p)])`` formula (numpy's
sequential ``cumsum``) wherever no comparison is within rounding of the threshold;
brute-force expansion sets; mass conservation and the sink budget of every snapshot.
"""
from __future__ import annotations

import numpy as np
from scipy.integrate import solve_ivp

from .fspmatrix import FspMatrix
from .statespace import StateSpace

RSTEP = "rstep"
SELECTIVE = "selective"
EPS = np.finfo(np.float64).eps  # Julia's eps()
TILE = 1024


def order_keys(p):
    """uint64 keys whose unsigned order is the IEEE total order of ``p`` (Julia ``isless``)."""
    u = np.ascontiguousarray(p, dtype=np.float64).view(np.uint64)
    neg = (u >> np.uint64(63)) == 1
    return np.where(neg, ~u, u | np.uint64(1 << 63))


def fixed_scan(x):
    """Inclusive prefix sum in the library's association order."""
    x = np.asarray(x, dtype=np.float64)
    n = x.size
    if n == 0:
        return x.copy()
    tiles = -(-n // TILE)
    a = np.zeros(tiles * TILE)
    a[:n] = x
    a = a.reshape(tiles, TILE)
    d = 1
    while d < TILE:
        a[:, d:] = a[:, d:] + a[:, :-d]  # right side is evaluated before the assignment
        d *= 2
    res = a.reshape(-1)[:n].copy()
    if tiles > 1:
        pre = fixed_scan(a[:, -1].copy())
        res[TILE:] = res[TILE:] + np.repeat(pre[:-1], TILE)[: n - TILE]
    return res


def drop_ids(p, t, tend, fsptol, strict):
    """Sorted 0-based ids that ``adapt!`` deletes (rstepadapters.jl:41-43 / 94-96)."""
    p = np.asarray(p, dtype=np.float64)
    if p.size == 0:
        return np.zeros(0, dtype=np.int64)
    perm = np.argsort(order_keys(p), kind="stable")
    cs = fixed_scan(p[perm])
    rest = cs[-1] - cs
    thr = 1.0 - t * fsptol / tend
    cnt = int(np.count_nonzero(rest > thr if strict else rest >= thr))
    return np.sort(perm[:cnt]).astype(np.int64)


def init(space: StateSpace, initial_step_count, p):
    """``init!`` (rstepadapters.jl:23-28, 74-79): returns p padded with zeros."""
    nold = space.count
    space.expand(initial_step_count)
    return np.concatenate([np.asarray(p, dtype=np.float64), np.zeros(space.count - nold)])


def adapt(space: StateSpace, kind, max_step_count, dropstates, p, t, tend, fsptol, dsinks=None):
    """``adapt!`` (rstepadapters.jl:35-52, 86-110). Returns ``(p_new, dropids0)``."""
    p = np.asarray(p, dtype=np.float64)
    if kind == SELECTIVE and p.size == 0:
        raise ValueError("Empty `p` input in `adapt!`.")
    ids = np.zeros(0, dtype=np.int64)
    if dropstates:
        ids = drop_ids(p, t, tend, fsptol, strict=(kind == SELECTIVE))
        space.deleteat(ids + 1)
        p = np.delete(p, ids)
    rs = ()
    if kind == SELECTIVE:
        rs = tuple(r + 1 for r in range(space.nr) if dsinks[r] > 0)
    nold = space.count
    space.expand(max_step_count, rs)
    return np.concatenate([p, np.zeros(space.count - nold)]), ids


def restart_sinks(sinks, t, tend, fsptol):
    """fspsolve.jl:178-181."""
    sinks = np.asarray(sinks, dtype=np.float64).copy()
    total = 0.0
    for s in sinks:  # sequential, as the library sums the R sink entries
        total += s
    if total >= t * fsptol / tend:
        sinks -= EPS
    return sinks


class Output:
    """Adaptive ``FspOutputSparse``: per snapshot ``t``, ``states``, ``p``, ``sinks``."""

    def __init__(self):
        self.t, self.states, self.p, self.sinks = [], [], [], []
        self.spaces = []  # state list of every segment
        self.events = []  # (t, u) at every budget event

    def push(self, t, states, u, S):
        self.t.append(float(t))
        self.states.append(states)
        self.p.append(u[:S].copy())
        self.sinks.append(u[S:].copy())

    def __len__(self):
        return len(self.t)


def solve_adaptive(stoich, propensities, parameters, init_states, init_values, tspan, kind, initial_step_count,
                   max_step_count, dropstates, saveat, fsptol, rtol=1e-12, atol=1e-15, max_adaptations=1000):
    """Adaptive FSP solve (fspsolve.jl:105-197) with a tight DOP853 integrator."""
    tstart, tend = min(tspan), max(tspan)
    saveat = np.asarray(saveat, dtype=np.float64)
    space = StateSpace(stoich, init_states)
    p0 = init(space, initial_step_count, init_values)
    R = space.nr
    u = np.concatenate([p0, np.zeros(R)])
    tnow = tstart
    out = Output()
    while tnow < tend:
        A = FspMatrix(space, propensities, parameters)
        S = space.count
        states = space.states.copy()
        out.spaces.append(states)

        def budget(t, y, S=S):
            return np.sum(y[S:]) - fsptol * t / tend

        budget.terminal = True
        budget.direction = 1.0
        for ts in saveat:
            if ts == tnow:
                out.push(tnow, states, u, S)
        te_pts = [float(ts) for ts in saveat if tnow < ts <= tend]
        eval_pts = sorted(set(te_pts + [tend]))
        sol = solve_ivp(lambda t, y: A.matvec(t, y), (tnow, tend), u, method="DOP853", rtol=rtol, atol=atol,
                        t_eval=eval_pts, events=budget)
        if sol.status not in (0, 1):
            raise RuntimeError(sol.message)
        reached = {float(t): k for k, t in enumerate(sol.t)}
        for ts in te_pts:
            if ts in reached:
                out.push(ts, states, sol.y[:, reached[ts]], S)
        if sol.status == 1:  # terminated by the budget event
            tnow = float(sol.t_events[0][0])
            u_ev = sol.y_events[0][0].copy()
            out.events.append((tnow, u_ev))
            if len(out.events) > max_adaptations:
                raise RuntimeError("too many adaptations")
            dsinks = A.matvec(tnow, u_ev)[S:] if kind == SELECTIVE else None
            p, _ = adapt(space, kind, max_step_count, dropstates, u_ev[:S], tnow, tend, fsptol, dsinks)
            u = np.concatenate([p, restart_sinks(u_ev[S:], tnow, tend, fsptol)])
        else:
            tnow = tend
            out.push(tend, states, sol.y[:, reached[tend]], S)
    return out
