"""GPU parity of the forward-sensitivity operator and solves.

* The fused (P+1)-block mat-vec against the oracle's ForwardSensFspMatrixSparse
  restatement (hand-derived propensity partials) on the telegraph, toggle (Hill terms,
  including the exponent as a parameter) and four-state gene (separable k21) models, and
  against the reference's analytic telegraph derivative matrices
  (test/sensmat/telegraph.jl:190-217, inf-norm <= n eps), plus its zero-sum tests.
* The sensitivity solve against the oracle's 13-block DOP853 solve (relative 1e-6), and
  the adaptive sensitivity solve of test/test_sensfsp.jl with invariants and bit-exact
  replays of every adaptation (fsspaceadapterssparse.jl:36-60).
"""
import numpy as np
import pytest

import models
import numcme_hip as nh
from oracle import adaptive as OA
from oracle.sensmatrix import SensFspMatrix, solve_sens_fixed
from oracle.statespace import StateSpace
from test_oracle_sens import _analytic

pytestmark = pytest.mark.gpu


def sens_model(mk, theta=None):
    S, props, th = mk() if theta is None else mk(theta)
    return S, props, th, nh.CmeModelWithSensitivity(nh.CmeModel(S, models.to_product(props), th))


CASES = {
    "telegraph_sens": (models.telegraph_sens, models.telegraph_sens_grads, models.telegraph_states(60)),
    "toggle": (models.toggle, models.toggle_grads, models.box_states([50, 50])),
    "gene4": (models.gene4, models.gene4_grads, models.gene4_states(8)),
}


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("case", list(CASES))
def test_sens_matvec_matches_oracle(gpu, case, precision):
    mk, gk, states = CASES[case]
    S, props, th, model = sens_model(mk)
    space = nh.StateSpaceSparse(S, states)
    SA = nh.ForwardSensFspMatrixSparse(model, space, precision=precision)
    ref = SensFspMatrix(StateSpace(S, states), models.to_oracle(props), th, gk())
    assert SA.rowcount == ref.n and SA.parameter_count == len(th)
    rng = np.random.default_rng(7)
    v = rng.random(ref.n * (len(th) + 1))
    for t in (0.0, 10.0, 77.0, 400.0):
        out = SA.matvec(t, v)
        vin = v if precision == "fp64" else v.astype(np.float32).astype(np.float64)
        want = ref.matvec(t, vin)
        scale = np.abs(want).max()
        tol = 1e-13 if precision == "fp64" else 2e-6
        assert np.abs(out - want).max() <= tol * scale, (case, t)


@pytest.mark.parametrize("t", [10.0, 20.0, 30.0, 100.0])
def test_sens_matvec_matches_reference_analytic(gpu, t):
    nmax = 500
    S, props, th, model = sens_model(models.telegraph_sens)
    space = nh.StateSpaceSparse(S, models.telegraph_states(nmax))
    SA = nh.ForwardSensFspMatrixSparse(model, space)
    n = SA.rowcount
    v = np.ones(n * 6)
    v /= v.sum()
    out = SA.matvec(t, v)
    A, dAs = _analytic(nmax, t, th)
    ref = np.zeros_like(v)
    ref[:n] = A @ v[:n]
    for ip in range(5):
        ref[(ip + 1) * n:(ip + 2) * n] = dAs[ip] @ v[:n] + A @ v[(ip + 1) * n:(ip + 2) * n]
    assert np.abs(out - ref).max() <= n * np.finfo(float).eps


def test_sens_matvec_zero_sums(gpu):
    S, props, th, model = sens_model(models.telegraph_sens)
    SA = nh.ForwardSensFspMatrixSparse(model, nh.StateSpaceSparse(S, models.telegraph_states(500)))
    v = np.ones(SA.rowcount * 6)
    out = SA.matvec(20.0, v)
    assert abs(out.sum() / v.sum()) <= SA.rowcount * 6 * np.finfo(float).eps
    Sp = np.array([[1, -1]])
    m = nh.CmeModelWithSensitivity(nh.CmeModel(Sp, [nh.propensity(lambda x, p: p[0]),
                                                    nh.propensity(lambda x, p: p[1])], [10.0, 5.0]))
    SA = nh.ForwardSensFspMatrixSparse(m, nh.StateSpaceSparse(Sp, [[i] for i in range(1, 101)]))
    out = SA.matvec(0.0, np.ones(3 * SA.rowcount))
    assert abs(out.sum()) <= 1e-9


def _fixed_ic(states, n_par):
    p0 = np.zeros(len(states))
    p0[0] = 1.0
    return nh.forwardsens_initial_condition(states, p0, [np.zeros(len(states)) for _ in range(n_par)])


# FP32-storage bars per case, just above the measured worst relative error
FP32_SENS_BAR = {"telegraph_sens": 1e-6, "gene4": 1e-5}


@pytest.mark.parametrize("case,ts,kinks", [
    ("telegraph_sens", [0.0, 5.0, 12.0, 25.0, 40.0], [10.0]),
    ("gene4", [0.0, 60.0, 120.0, 300.0], []),
])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_sens_solve_matches_oracle(gpu, case, ts, kinks, precision):
    mk, gk, _ = CASES[case]
    states = models.telegraph_states(40) if case == "telegraph_sens" else models.gene4_states(5)
    S, props, th, model = sens_model(mk)
    ic = _fixed_ic(states, len(th))
    ts = np.asarray(ts)
    # the fixed-space form: no initial expansion and a budget the sinks never reach
    alg = nh.AdaptiveForwardSensFspSparse(space_adapter=nh.ForwardSensRStepAdapter(0, 1, False),
                                          ode_method=nh.RK45(rtol=1e-10, atol=1e-16, precision=precision))
    out = nh.solve(model, ic, (ts[0], ts[-1]), alg, saveat=ts, fsptol=1.0)
    ref = solve_sens_fixed(StateSpace(S, states), models.to_oracle(props), th, gk(), ic.p, ic.S, ts,
                           breakpoints=kinks)
    assert out.stats["adaptations"] == 0 and len(out) == len(ts) + 1
    # relative to each sensitivity's largest entry; FP32 storage measures 6.5e-7 on the
    # full-size C3 solve
    bar = 1e-6 if precision == "fp64" else FP32_SENS_BAR[case]
    worst, worst_sink = 0.0, 0.0
    for k in range(len(ts)):
        assert np.array_equal(out.p[k].states, ref.states)
        assert np.abs(out.p[k].values - ref.p[k]).sum() <= (1e-8 if precision == "fp64" else 1e-6)
        for ip in range(len(th)):
            want = ref.S[k][ip]
            scale = max(np.abs(want).max(), 1e-12)
            worst = max(worst, np.abs(out.S[k][ip].values - want).max() / scale)
            worst_sink = max(worst_sink, (np.abs(out.dsinks[k][ip] - ref.dsinks[k][ip]).max() - 1e-14) / scale)
    print(f"{case} {precision}: worst relative sensitivity error {worst:.2e}, sinks {worst_sink:.2e}", flush=True)
    assert worst <= bar and worst_sink <= bar


def test_reference_adaptive_sens_solve(gpu):
    # test/test_sensfsp.jl: ForwardSensRStepAdapter(10, 10, true) over (0, 400)
    S, props, th, model = sens_model(models.telegraph_sens)
    ic = nh.forwardsens_initial_condition([[1, 0, 0]], [1.0], [[0.0] for _ in range(5)])
    alg = nh.AdaptiveForwardSensFspSparse(space_adapter=nh.ForwardSensRStepAdapter(10, 10, True),
                                          ode_method=nh.RK45(rtol=1e-8, atol=1e-14))
    out = nh.solve(model, ic, (0.0, 400.0), alg, saveat=[], fsptol=1e-6, odeatol=1e-10, odertol=1e-4)
    assert out.stats["adaptations"] >= 1
    assert out.t[-1] == out.t[-2] == 400.0
    sid = np.asarray(out.space_ids)
    for k in range(len(out)):
        assert out.sinks[k].sum() <= 1e-6 * out.t[k] / 400.0 + 1e-13
        for ip in range(5):
            # the total mass does not depend on theta, up to the mass the drops remove
            assert abs(out.S[k][ip].values.sum() + out.dsinks[k][ip].sum()) <= 1e-6
    # every adaptation replayed by the oracle from the library's own event state
    for s in range(sid[-1]):
        last = int(np.nonzero(sid == s)[0][-1])
        first = last + 1
        t_ev = out.t[last]
        osp = StateSpace(S, out.p[last].states)
        p_new, ids = OA.adapt(osp, OA.RSTEP, 10, True, out.p[last].values, t_ev, 400.0, 1e-6)
        assert np.array_equal(out.p[first].states, osp.states)
        assert np.array_equal(out.p[first].values, p_new)
        keep = np.ones(len(out.p[last].values), dtype=bool)
        keep[ids] = False
        for ip in range(5):
            s_new = np.concatenate([out.S[last][ip].values[keep], np.zeros(osp.count - keep.sum())])
            assert np.array_equal(out.S[first][ip].values, s_new)
            assert np.array_equal(out.dsinks[first][ip], out.dsinks[last][ip])
        assert np.array_equal(out.sinks[first], OA.restart_sinks(out.sinks[last], t_ev, 400.0, 1e-6))


@pytest.mark.parametrize("order", ["lattice", "shuffled"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_sens_matvec_twelve_slots(gpu, order, precision):
    # 12 slots: the generic (runtime slot count) sensitivity kernel, columns computed
    # (lattice order) or loaded (shuffled)
    S, props, th = models.twelve_slot()
    model = nh.CmeModelWithSensitivity(nh.CmeModel(S, models.to_product(props), th))
    states = models.box_states([6, 5, 4])
    if order == "shuffled":
        states = states[np.random.default_rng(8).permutation(len(states))]
    SA = nh.ForwardSensFspMatrixSparse(model, nh.StateSpaceSparse(S, states), precision=precision)
    ref = SensFspMatrix(StateSpace(S, states), models.to_oracle(props), th, models.twelve_slot_grads())
    v = np.random.default_rng(9).random(ref.n * (len(th) + 1))
    out = SA.matvec(0.0, v)
    vin = v if precision == "fp64" else v.astype(np.float32).astype(np.float64)
    want = ref.matvec(0.0, vin)
    tol = 1e-13 if precision == "fp64" else 2e-6
    assert np.abs(out - want).max() <= tol * np.abs(want).max()


def _birth_death(n_par):
    # one species, birth k and death gamma x; with one parameter gamma is fixed at 0.5
    S = np.array([[1, -1]])
    if n_par == 2:
        props = [("ti", lambda x, p: p[0] + 0.0 * x[0]), ("ti", lambda x, p: p[1] * x[0])]
        th = np.array([12.0, 0.7])
    else:
        props = [("ti", lambda x, p: p[0] + 0.0 * x[0]), ("ti", lambda x, p: 0.5 * x[0])]
        th = np.array([12.0])
    return S, props, th


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("n_par", [1, 2])
def test_sens_matvec_few_blocks(gpu, n_par, precision):
    # two or three vector blocks: some of the four waves of a workgroup own no block
    from oracle.sensmatrix import DTimeInvariant
    S, props, th = _birth_death(n_par)
    grads = {(0, 0): DTimeInvariant(lambda x, p: 1.0 + 0.0 * x[0])}
    if n_par == 2:
        grads[(1, 1)] = DTimeInvariant(lambda x, p: 1.0 * x[0])
    states = [[i] for i in range(0, 300)]
    model = nh.CmeModelWithSensitivity(nh.CmeModel(S, models.to_product(props), th))
    SA = nh.ForwardSensFspMatrixSparse(model, nh.StateSpaceSparse(S, states), precision=precision)
    ref = SensFspMatrix(StateSpace(S, states), models.to_oracle(props), th, grads)
    assert SA.rowcount == ref.n and SA.parameter_count == n_par
    v = np.random.default_rng(3).random(ref.n * (n_par + 1))
    out = SA.matvec(0.0, v)
    vin = v if precision == "fp64" else v.astype(np.float32).astype(np.float64)
    want = ref.matvec(0.0, vin)
    tol = 1e-13 if precision == "fp64" else 2e-6
    assert np.abs(out - want).max() <= tol * np.abs(want).max()
