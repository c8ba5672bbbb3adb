"""Reaction networks of the test suite, written once and handed to both the oracle and
the GPU library.

Propensities use ``numcme_hip.expr`` functions, which trace symbolically for the device
and evaluate with numpy for the oracle. Species and parameters are 0-based.
"""
from __future__ import annotations

import numpy as np

from numcme_hip import expr as nm

# ---- telegraph gene model (test/test_fspmat.jl, test/test_solver.jl, config C1) ----------
TELEGRAPH_S = np.array([[-1, 1, 0, 0], [1, -1, 0, 0], [0, 0, 1, -1]])


def telegraph(theta=(0.05, 0.1, 5.0, 0.05)):
    """G_off->G_on k_on, G_on->G_off k_off, G_on->G_on+mRNA k_r, mRNA->0 gamma."""
    props = [
        ("ti", lambda x, p: p[0] * x[0]),
        ("ti", lambda x, p: p[1] * x[1]),
        ("ti", lambda x, p: p[2] * x[1]),
        ("ti", lambda x, p: p[3] * x[2]),
    ]
    return TELEGRAPH_S, props, np.asarray(theta, dtype=float)


def telegraph_tv(kind, theta=(0.05, 0.1, 5.0, 1.0)):
    """test/test_fspmat.jl:103-124: reaction 2 scaled by max(0, 1 - sin(pi t / 2))."""
    tf = lambda t, p: nm.max(0.0, 1.0 - nm.sin(nm.pi * t / 2))
    a2 = ("sep", tf, lambda x, p: p[1] * x[1]) if kind == "sep" else \
         ("joint", lambda t, x, p: nm.max(0.0, 1.0 - nm.sin(nm.pi * t / 2)) * p[1] * x[1])
    props = [("ti", lambda x, p: p[0] * x[0]), a2, ("ti", lambda x, p: p[2] * x[1]), ("ti", lambda x, p: p[3] * x[2])]
    return TELEGRAPH_S, props, np.asarray(theta, dtype=float)


def telegraph_sens(theta=(0.05, 0.1, 5.0, 0.5, 20.0)):
    """test/sensmat/telegraph.jl:7-28: degradation scaled by max(0, 1 - sin(pi t / L))."""
    props = [
        ("ti", lambda x, p: p[0] * x[0]),
        ("ti", lambda x, p: p[1] * x[1]),
        ("ti", lambda x, p: p[2] * x[1]),
        ("sep", lambda t, p: nm.max(0.0, 1.0 - nm.sin(nm.pi * t / p[4])), lambda x, p: p[3] * x[2]),
    ]
    return TELEGRAPH_S, props, np.asarray(theta, dtype=float)


def telegraph_states(nmax):
    """Gene-major rectangular telegraph space (the convention of test/sensmat/telegraph.jl:31)."""
    return [[1, 0, m] for m in range(nmax + 1)] + [[0, 1, m] for m in range(nmax + 1)]


# ---- toggle switch (config C2) -----------------------------------------------------------
TOGGLE_S = np.array([[1, -1, 0, 0], [0, 0, 1, -1]])


def toggle(theta=(50.0, 20.0, 2.5, 1.0, 1.0)):
    """Production of S1: b + alpha/(1+(S2/K)^n); first-order degradation gamma."""
    props = [
        ("ti", lambda x, p: p[3] + p[0] / (1.0 + (x[1] / p[1]) ** p[2])),
        ("ti", lambda x, p: p[4] * x[0]),
        ("ti", lambda x, p: p[3] + p[0] / (1.0 + (x[0] / p[1]) ** p[2])),
        ("ti", lambda x, p: p[4] * x[1]),
    ]
    return TOGGLE_S, props, np.asarray(theta, dtype=float)


def box_states(bounds):
    """All states 0 <= x_s <= bounds[s], species 0 varying fastest."""
    grids = np.meshgrid(*[np.arange(b + 1) for b in bounds[::-1]], indexing="ij")
    return np.stack([g.ravel() for g in grids[::-1]], axis=1).astype(np.int64)


# ---- four-state gene, nuclear and cytoplasmic mRNA (config C3) ---------------------------
def gene4_stoich():
    S = np.zeros((6, 12), dtype=np.int64)
    moves = [(0, 1), (1, 0), (1, 2), (2, 1), (2, 3), (3, 2)]
    for r, (a, b) in enumerate(moves):
        S[a, r] -= 1
        S[b, r] += 1
    for r in range(6, 10):
        S[4, r] = 1
    S[4, 10], S[5, 10] = -1, 1
    S[5, 11] = -1
    return S


GENE4_THETA = (1.3e-3, 3.2e-3, 5.9e-3, 2e-2, 1.5e-3, 4.6e-3, 0.0, 1e-2, 0.2, 1.5, 0.05, 1e-3)


def gene4(theta=GENE4_THETA):
    """Gene states 0..3 (one-hot), nuclear (4) and cytoplasmic (5) mRNA; k21 pulse-modulated."""
    sig = lambda t: (1.0 - nm.exp(-6e-3 * t)) * nm.exp(-9e-3 * t)
    props = [
        ("ti", lambda x, p: p[0] * x[0]),
        ("ti", lambda x, p: p[1] * x[1]),
        ("ti", lambda x, p: p[2] * x[1]),
        ("sep", lambda t, p: nm.max(0.0, p[3] - 0.12 * sig(t)), lambda x, p: x[2]),
        ("ti", lambda x, p: p[4] * x[2]),
        ("ti", lambda x, p: p[5] * x[3]),
        ("ti", lambda x, p: p[6] * x[0]),
        ("ti", lambda x, p: p[7] * x[1]),
        ("ti", lambda x, p: p[8] * x[2]),
        ("ti", lambda x, p: p[9] * x[3]),
        ("ti", lambda x, p: p[10] * x[4]),
        ("ti", lambda x, p: p[11] * x[5]),
    ]
    return gene4_stoich(), props, np.asarray(theta, dtype=float)


def gene4_states(nmax):
    out = []
    for g in range(4):
        oh = [0, 0, 0, 0]
        oh[g] = 1
        for c in range(nmax + 1):
            for n in range(nmax + 1):
                out.append(oh + [n, c])
    return np.asarray(out, dtype=np.int64)


# ---- adapters ------------------------------------------------------------------------------
def to_oracle(props):
    from oracle.fspmatrix import Joint, Separable, TimeInvariant
    out = []
    for pr in props:
        if pr[0] == "ti":
            out.append(TimeInvariant(pr[1]))
        elif pr[0] == "sep":
            out.append(Separable(pr[1], pr[2]))
        else:
            out.append(Joint(pr[1]))
    return out


def to_product(props):
    from numcme_hip import (JointTimeVaryingPropensity, SeparableTimeVaryingPropensity,
                            StandardTimeInvariantPropensity)
    out = []
    for pr in props:
        if pr[0] == "ti":
            out.append(StandardTimeInvariantPropensity(pr[1]))
        elif pr[0] == "sep":
            out.append(SeparableTimeVaryingPropensity(pr[1], pr[2]))
        else:
            out.append(JointTimeVaryingPropensity(pr[1]))
    return out


# ---- three-species cascade (config C5) ----------------------------------------------------
CASCADE_S = np.array([[1, 0, 0, -1, 0, 0], [0, 1, 0, 0, -1, 0], [0, 0, 1, 0, 0, -1]])


def cascade(theta=(10.0, 2.0, 1.0, 0.5, 0.2, 0.1)):
    """0->A k0; A->A+B k1*A; B->B+C k2*B; A->0, B->0, C->0 first order."""
    props = [
        ("ti", lambda x, p: p[0] + 0.0 * x[0]),
        ("ti", lambda x, p: p[1] * x[0]),
        ("ti", lambda x, p: p[2] * x[1]),
        ("ti", lambda x, p: p[3] * x[0]),
        ("ti", lambda x, p: p[4] * x[1]),
        ("ti", lambda x, p: p[5] * x[2]),
    ]
    return CASCADE_S, props, np.asarray(theta, dtype=float)


# ---- hand-derived propensity partial derivatives (the oracle's gradient input) ----------
# Keys (r, p) form the gradient sparsity pattern. Julia's max(0, y) is y unless y < 0, so
# its derivative is dy where y >= 0 and 0 elsewhere.
def _dmax0(y, dy):
    return np.where(np.asarray(y) >= 0, dy, 0.0)


def telegraph_sens_grads():
    from oracle.sensmatrix import DSeparable, DTimeInvariant
    tf_arg = lambda t, p: np.pi * t / p[4]
    return {
        (0, 0): DTimeInvariant(lambda x, p: 1.0 * x[0]),
        (1, 1): DTimeInvariant(lambda x, p: 1.0 * x[1]),
        (2, 2): DTimeInvariant(lambda x, p: 1.0 * x[1]),
        (3, 3): DSeparable(None, lambda x, p: 1.0 * x[2]),
        (3, 4): DSeparable(lambda t, p: _dmax0(1.0 - np.sin(tf_arg(t, p)), np.cos(tf_arg(t, p)) * np.pi * t / p[4] ** 2),
                           None),
    }


def gene4_grads():
    from oracle.sensmatrix import DSeparable, DTimeInvariant
    sig = lambda t: (1.0 - np.exp(-6e-3 * t)) * np.exp(-9e-3 * t)
    species = {0: 0, 1: 1, 2: 1, 4: 2, 5: 3, 6: 0, 7: 1, 8: 2, 9: 3, 10: 4, 11: 5}
    g = {(r, r): DTimeInvariant((lambda s: lambda x, p: 1.0 * x[s])(s)) for r, s in species.items()}
    g[(3, 3)] = DSeparable(lambda t, p: _dmax0(p[3] - 0.12 * sig(t), 1.0), None)
    return g


def toggle_grads():
    from oracle.sensmatrix import DTimeInvariant

    def hill_terms(y, p):
        y = np.asarray(y, dtype=np.float64)
        u = y / p[1]
        un = np.power(u, p[2])
        den = 1.0 + un
        with np.errstate(divide="ignore", invalid="ignore"):
            logu = np.where(u > 0, np.log(np.where(u > 0, u, 1.0)), 0.0)
        return un, den, logu

    def d_alpha(s):
        return lambda x, p: 1.0 / hill_terms(x[s], p)[1]

    def d_K(s):
        def f(x, p):
            un, den, _ = hill_terms(x[s], p)
            return p[0] * un * p[2] / p[1] / den ** 2
        return f

    def d_n(s):
        def f(x, p):
            un, den, logu = hill_terms(x[s], p)
            return -p[0] * un * logu / den ** 2
        return f

    g = {}
    for r, other in ((0, 1), (2, 0)):
        g[(r, 0)] = DTimeInvariant(d_alpha(other))
        g[(r, 1)] = DTimeInvariant(d_K(other))
        g[(r, 2)] = DTimeInvariant(d_n(other))
        g[(r, 3)] = DTimeInvariant(lambda x, p: 1.0 + 0.0 * x[0])
    g[(1, 4)] = DTimeInvariant(lambda x, p: 1.0 * x[0])
    g[(3, 4)] = DTimeInvariant(lambda x, p: 1.0 * x[1])
    return g


# ---- twelve distinct stoichiometries (slot counts above the fixed-count kernels) ---------
def twelve_slot():
    """Three species with births, deaths and all six conversions: 12 reactions with distinct
    stoichiometries, so a generator on a box has 12 slots."""
    cols, props = [], []
    for s in range(3):
        e = np.zeros(3, dtype=np.int64)
        e[s] = 1
        cols.append(e.copy())
        props.append(("ti", (lambda s: lambda x, p: p[0] * (1.0 + 0.1 * s))(s)))
        cols.append(-e)
        props.append(("ti", (lambda s: lambda x, p: p[1] * x[s])(s)))
    for a in range(3):
        for b in range(3):
            if a == b:
                continue
            c = np.zeros(3, dtype=np.int64)
            c[a], c[b] = -1, 1
            cols.append(c)
            props.append(("ti", (lambda a, b: lambda x, p: p[2] * (1 + a + 2 * b) * x[a])(a, b)))
    return np.stack(cols, axis=1), props, np.array([2.0, 0.3, 0.05])


def twelve_slot_grads():
    """Propensity partials of twelve_slot(), keyed (reaction, parameter): births
    p0 (1 + 0.1 s), deaths p1 x_s, conversions p2 (1 + a + 2 b) x_a."""
    from oracle.sensmatrix import DTimeInvariant
    g, r = {}, 0
    for s in range(3):
        g[(r, 0)] = DTimeInvariant((lambda s: lambda x, p: (1.0 + 0.1 * s) + 0.0 * x[0])(s))
        g[(r + 1, 1)] = DTimeInvariant((lambda s: lambda x, p: 1.0 * x[s])(s))
        r += 2
    for a in range(3):
        for b in range(3):
            if a != b:
                g[(r, 2)] = DTimeInvariant((lambda a, b: lambda x, p: (1.0 + a + 2 * b) * x[a])(a, b))
                r += 1
    return g
