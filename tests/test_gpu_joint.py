"""GPU parity of time-varying generators at generic times, against the oracle.

The joint refresh (``_update_sparsematrix!``, fspsparsematrix.jl:154-166) runs whenever
``matvec!`` sees a new ``t`` (:206-213). These tests evaluate the device generator at
t = 0.3, 0.3, 1.45, 0.3, 2.6 (the repeated 0.3 takes the cached-time path, the returns to
0.3 catch a refresh that keeps a stale time), where max(0, 1 - sin(pi t / 2)) is 0.546,
0.240 and 1.809: neither 0 nor 1, and distinct (t = 1.7 would repeat the value at 0.3).
Both FP64 and FP32, ``matvec!`` and ``matvecadd!``, sink rows included.
The oracle's joint matrix is itself pinned by the separable form (``test_oracle_fspmatrix``).

Three lowerings are covered: a factorable joint law kept joint (``factor_joint=False``,
the device refresh kernels), the same law factored to separable form (the default), and a
law that does not factor (x / (1 + t x)), which stays joint. RK45 solves of the joint
telegraph and of the non-factorable model are compared with the oracle's DOP853 solve:
l1 <= 1e-8 and sink mass <= 1e-10 per snapshot (FP64), l1 <= 1e-6 (FP32).
"""
import numpy as np
import pytest

import models
import numcme_hip as nh
from numcme_hip import _lib as L
from oracle.fspmatrix import FspMatrix
from oracle.solve import solve_fixed
from oracle.statespace import StateSpace

pytestmark = pytest.mark.gpu

TIMES = (0.3, 0.3, 1.45, 0.3, 2.6)


def nonfactorable(theta=(0.05, 0.1, 5.0, 1.0)):
    """Telegraph whose mRNA decay saturates in time and count: gamma x / (1 + 0.2 t x)."""
    props = [("ti", lambda x, p: p[0] * x[0]), ("ti", lambda x, p: p[1] * x[1]), ("ti", lambda x, p: p[2] * x[1]),
             ("joint", lambda t, x, p: p[3] * x[2] / (1.0 + 0.2 * t * x[2]))]
    return models.TELEGRAPH_S, props, np.asarray(theta, dtype=float)


def build(props, th, levels, precision, factor_joint):
    """Space from (1, 0, 0) expanded by ``levels``, or, for ``levels=None``, the gene-major
    box of 6,000 states (mRNA 0..2999; several workgroups per launch)."""
    S = models.TELEGRAPH_S
    init = [[1, 0, 0]] if levels is not None else models.telegraph_states(2999)
    d = nh.StateSpaceSparse(S, init)
    o = StateSpace(S, init)
    if levels is not None:
        d.expand(levels)
        o.expand(levels)
    assert np.array_equal(d.get_states(), o.states)
    assert np.array_equal(d.get_state_connectivity(), o.state_conn)
    model = nh.CmeModel(S, models.to_product(props), th, factor_joint=factor_joint)
    A = nh.FspMatrixSparse(d, None, precision=precision, model=model)
    return A, FspMatrix(o, models.to_oracle(props), th), model


CASES = [("telegraph_joint", False, L.NC_PROP_JOINT), ("telegraph_joint", True, L.NC_PROP_SEPARABLE),
         ("nonfactorable", True, L.NC_PROP_JOINT)]


def case_model(name, k_off=0.1):
    if name == "telegraph_joint":
        return models.telegraph_tv("joint", theta=(0.05, k_off, 5.0, 1.0))
    return nonfactorable()


@pytest.mark.parametrize("levels", [25, None])
@pytest.mark.parametrize("name,factor,lowered", CASES)
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_joint_matvec_generic_times(gpu, name, factor, lowered, precision, levels):
    # k_off = 40: the time-varying reaction's entries are as large as the mRNA terms, so
    # that its change with t stands far above the FP32 tolerance band
    S, props, th = case_model(name, k_off=40.0)
    A, Ao, model = build(props, th, levels, precision, factor)
    r_tv = 1 if name == "telegraph_joint" else 3
    assert model.programs.kinds[r_tv] == lowered
    n = A.rowcount
    n_states = n - 4
    rng = np.random.default_rng(7)
    v = rng.random(n)
    vin = v if precision == "fp64" else v.astype(np.float32).astype(np.float64)
    tol = 1e-13 if precision == "fp64" else 2e-6
    prev = None
    for t in TIMES:
        wo = Ao.matvec(t, vin)
        w = nh.matvec(t, A, v)
        scale = np.abs(wo).max()
        assert np.abs(w - wo).max() <= tol * scale, t
        # the sink rows, entry by entry
        assert np.allclose(w[n_states:], wo[n_states:], rtol=tol, atol=tol * scale)
        # matvecadd! adds into the output
        base = rng.random(n)
        acc = base.copy()
        A.matvecadd_(acc, t, v)
        assert np.abs((acc - base) - wo).max() <= tol * scale + 4e-16 * np.abs(base).max()
        # every column still sums to zero (mass is conserved at every t): the sum of A 1 is
        # zero up to the rounding of n row sums of entries as large as the diagonal
        dmax = np.abs(Ao.to_csc(t).diagonal()).max()
        assert abs(nh.matvec(t, A, np.ones(n)).sum()) <= 1e-15 * n * dmax
        if prev is not None and prev[0] != t:
            # the generator changes between these times by far more than the tolerance, so
            # a refresh that ignored t or kept a stale one fails the comparison above
            assert np.abs(wo - prev[1]).max() > 10 * 2e-6 * scale
        prev = (t, wo)


@pytest.mark.parametrize("name,factor,lowered", CASES)
def test_joint_export_matches_oracle(gpu, name, factor, lowered):
    S, props, th = case_model(name)
    A, Ao, _ = build(props, th, 12, "fp64", factor)
    for t in TIMES:
        D = (A.to_scipy(t) - Ao.to_csc(t)).tocoo()
        assert np.abs(D.data).max(initial=0.0) <= 1e-14 * abs(Ao.to_csc(t)).max(), t


@pytest.mark.parametrize("name,factor,lowered", CASES)
@pytest.mark.parametrize("precision,bar", [("fp64", 1e-8), ("fp32", 1e-6)])
def test_joint_rk45_solve_vs_oracle(gpu, name, factor, lowered, precision, bar):
    S, props, th = case_model(name)
    o = StateSpace(S, [[1, 0, 0]])
    o.expand(14)
    states = o.states
    p0 = np.zeros(len(states))
    p0[0] = 1.0
    ts = np.linspace(0.0, 4.0, 9)
    # the time factor of the telegraph law touches zero at t = 1 (1 - sin(pi/2) = 0): a
    # break point for the oracle's integrator
    ref = solve_fixed(S, models.to_oracle(props), th, states, p0, ts, breakpoints=(1.0, 3.0))
    model = nh.CmeModel(S, models.to_product(props), th, factor_joint=factor)
    out = nh.solve(model, nh.FspVectorSparse(states, p0), (0.0, 4.0), nh.RK45(rtol=1e-10, atol=1e-16, precision=precision),
                   saveat=ts)
    for k in range(len(ts)):
        assert np.array_equal(out.p[k].states, states)
        err = float(np.abs(out.p[k].values - ref.p[k]).sum())
        assert err <= bar, (k, err)
        if precision == "fp64":
            assert np.abs(out.sinks[k] - ref.sinks[k]).max() <= 1e-10
    # the sink of mRNA production holds mass: the boundary is reached
    assert ref.sinks[-1][2] > 1e-6
