"""GPU parity of the adaptive FSP solve and the R-step adapters.

* Drop selection and adapt! through the C ABI against the oracle, bit for bit (ids,
  state lists, probabilities), on up to 3.1M states, with ties, zeros, -0.0 and tiny
  negatives (rstepadapters.jl:41-43, 94-96).
* Every adaptation of a solve replayed by the oracle from the library's own event state:
  the next state list, the restart probabilities and the restart sinks must match bit for
  bit (rstepadapters.jl:35-110, fspsolve.jl:171-184).
* End-to-end adaptive solves without dropping against the oracle's exact adaptive solve
  (same state lists, probabilities within l1 1e-8), and the reference's own adaptive test
  (test/test_solver.jl:79-100: mass conservation, parameter-free vs parametric l1 <= 1e-14).
"""
import ctypes as C

import numpy as np
import pytest

import models
from fspcompare import l1_distance
import numcme_hip as nh
from numcme_hip import _lib as L
from oracle import adaptive as OA
from oracle.fspmatrix import FspMatrix
from oracle.statespace import StateSpace

pytestmark = pytest.mark.gpu


def l1(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).sum())


def random_p(n, seed):
    rng = np.random.default_rng(seed)
    p = rng.exponential(size=n) ** 4
    p[rng.random(n) < 0.1] = 0.0
    p[rng.random(n) < 0.02] = -0.0
    p[rng.random(n) < 0.01] = -1e-18 * rng.random()
    dup = rng.integers(0, n, size=max(1, n // 20))
    p[dup] = p[dup[0]]  # a block of exact ties
    return p / np.abs(p).sum()


@pytest.mark.parametrize("n_side,levels", [(3, 1), (40, 2), (550, 1), (1024, 0), (1765, 0)])
@pytest.mark.parametrize("kind", [OA.RSTEP, OA.SELECTIVE])
def test_adapt_matches_oracle(gpu, n_side, levels, kind):
    S, _, _ = models.toggle()
    box = models.box_states([n_side - 1, n_side - 1])
    n = len(box)
    p = random_p(n, n_side)
    fsptol, t, tend = 0.02, 1.0, 1.0
    ad = (nh.SelectiveRStepAdapter if kind == OA.SELECTIVE else nh.RStepAdapter)(0, levels, True)
    dsinks = np.array([1.0, 0.0, 2.0, 0.0])
    print(f"adapt n={n} kind={kind}", flush=True)
    sp = nh.StateSpaceSparse(S, box)
    p_new, drop1 = nh.adapt_(sp, ad, p, None, t, tend, fsptol, dsinks=dsinks)
    ref_ids = OA.drop_ids(p, t, tend, fsptol, strict=(kind == OA.SELECTIVE))
    assert ref_ids.size > 0 or n < 100
    assert np.array_equal(drop1 - 1, ref_ids)
    if n <= 400_000:
        osp = StateSpace(S, box)
        p_ref, ids = OA.adapt(osp, kind, levels, True, p, t, tend, fsptol, dsinks)
        assert np.array_equal(ids, ref_ids)
        assert np.array_equal(sp.get_states(), osp.states)
        assert np.array_equal(p_new, p_ref)
    else:
        assert sp.get_state_count() == n - ref_ids.size


def test_adapt_exact_ties_separate_ge_from_gt(gpu):
    # dyadic probabilities add exactly: the threshold meets sum - cumsum exactly, so RStep
    # (>=) drops one state more than SelectiveRStep (>)
    S, _, _ = models.toggle()
    box = models.box_states([63, 63])
    rng = np.random.default_rng(5)
    w = rng.integers(0, 16, size=len(box)).astype(np.float64)
    p = w / 2.0**15
    total = p.sum()
    perm = np.argsort(OA.order_keys(p), kind="stable")
    rest = total - np.cumsum(p[perm])
    k = int(np.nonzero(rest < total - 0.2)[0][0])
    q = total - rest[k]  # the mass of the k+1 smallest states, dyadic
    fsptol = (1.0 - total) + q  # thr = 1 - fsptol = total - q = rest[k]
    counts = {}
    for kind, cls in [(OA.RSTEP, nh.RStepAdapter), (OA.SELECTIVE, nh.SelectiveRStepAdapter)]:
        ref = OA.drop_ids(p, 1.0, 1.0, fsptol, strict=(kind == OA.SELECTIVE))
        sp = nh.StateSpaceSparse(S, box)
        _, d = nh.adapt_(sp, cls(0, 0, True), p, None, 1.0, 1.0, fsptol, dsinks=np.ones(4))
        assert np.array_equal(d - 1, ref)
        counts[kind] = d.size
    assert counts[OA.RSTEP] > counts[OA.SELECTIVE]


def test_adapt_errors(gpu):
    S, _, _ = models.telegraph()
    sp = nh.StateSpaceSparse(S, [[1, 0, 0]])
    sp.deleteat([1])
    with pytest.raises(nh.ArgumentError):
        nh.adapt_(sp, nh.SelectiveRStepAdapter(1, 1, True), np.zeros(0), None, 1.0, 2.0, 1e-4, dsinks=np.zeros(4))
    # the C entry point itself rejects the empty p (rstepadapters.jl:90-92)
    ad = L.nc_adapter(L.NC_ADAPTER_SELECTIVE_RSTEP, 1, 1, 1)
    m = C.c_int64()
    z = np.zeros(4)
    rc = L.lib().nc_adapter_adapt(sp.handle, C.byref(ad), 0, L.ptr(z, C.c_double), L.ptr(z, C.c_double), 1.0, 2.0,
                                  1e-4, C.byref(m), None)
    assert rc == L.NC_ERR_ARG
    sp2 = nh.StateSpaceSparse(S, [[1, 0, 0]])
    with pytest.raises(nh.ArgumentError):
        nh.adapt_(sp2, nh.RStepAdapter(0, 1, True), np.ones(3), None, 1.0, 2.0, 1e-4)  # wrong length


def test_init_expands_all_reactions(gpu):
    S, _, _ = models.toggle()
    sp = nh.StateSpaceSparse(S, [[0, 0]])
    p = nh.init_(sp, nh.RStepAdapter(3, 1, False), [1.0])
    osp = StateSpace(S, [[0, 0]])
    pr = OA.init(osp, 3, [1.0])
    assert np.array_equal(sp.get_states(), osp.states) and np.array_equal(p, pr)


def _telegraph_model(theta=(0.05, 0.1, 5.0, 1.0)):
    S, props, th = models.telegraph(theta)
    return S, props, th, nh.CmeModel(S, models.to_product(props), th)


@pytest.mark.parametrize("method", ["krylov", "rk45"])
def test_adaptive_rstep_matches_oracle(gpu, method):
    S, props, th, model = _telegraph_model()
    ts = np.arange(0.0, 121.0, 20.0)
    fsptol = 1e-4
    ref = OA.solve_adaptive(S, models.to_oracle(props), th, [[1, 0, 0]], [1.0], (0.0, 120.0), OA.RSTEP, 3, 3, False,
                            ts, fsptol)
    ode = nh.KrylovExpv(m=30, tol=1e-13) if method == "krylov" else nh.RK45(rtol=1e-11, atol=1e-16)
    alg = nh.AdaptiveFspSparse(ode_method=ode, space_adapter=nh.RStepAdapter(3, 3, False))
    out = nh.solve(model, nh.FspVectorSparse([[1, 0, 0]], [1.0]), (0.0, 120.0), alg, saveat=ts, fsptol=fsptol)
    assert len(out) == len(ref) == len(ts) + 1
    assert out.stats["adaptations"] == len(ref.events)
    for k in range(len(out)):
        assert out.t[k] == ref.t[k]
        assert np.array_equal(out.p[k].states, ref.states[k])
        assert l1(out.p[k].values, ref.p[k]) <= 1e-8
        assert l1(out.sinks[k], ref.sinks[k]) <= 1e-10


@pytest.mark.parametrize("kind", [OA.RSTEP, OA.SELECTIVE])
@pytest.mark.parametrize("method", ["krylov", "rk45"])
def test_adaptations_replay_bit_exact(gpu, kind, method):
    # empty saveat stores every step, so the event state (last snapshot of a space) and the
    # restart state (first snapshot of the next) are both in the output
    S, props, th, model = _telegraph_model()
    fsptol, tend = 1e-4, 60.0
    cls = nh.SelectiveRStepAdapter if kind == OA.SELECTIVE else nh.RStepAdapter
    ode = nh.KrylovExpv(m=30, tol=1e-12) if method == "krylov" else nh.RK45(rtol=1e-9, atol=1e-14)
    alg = nh.AdaptiveFspSparse(ode_method=ode, space_adapter=cls(2, 2, True))
    out = nh.solve(model, nh.FspVectorSparse([[1, 0, 0]], [1.0]), (0.0, tend), alg, fsptol=fsptol)
    sid = np.asarray(out.space_ids)
    assert out.stats["adaptations"] >= 3 and sid[-1] == out.stats["adaptations"]
    assert out.t[-1] == out.t[-2] == tend
    oprops = models.to_oracle(props)
    dropped = 0
    for s in range(sid[-1]):
        last = int(np.nonzero(sid == s)[0][-1])
        first = last + 1
        assert sid[first] == s + 1 and out.t[first] == out.t[last]
        t_ev = out.t[last]
        p_ev, sinks_ev = out.p[last].values, out.sinks[last]
        # the event: the budget is crossed at t_ev
        assert sinks_ev.sum() >= fsptol * t_ev / tend - 1e-10
        osp = StateSpace(S, out.p[last].states)
        dsinks = None
        if kind == OA.SELECTIVE:
            dsinks = FspMatrix(osp, oprops, th).matvec(t_ev, np.concatenate([p_ev, sinks_ev]))[osp.count:]
        p_new, ids = OA.adapt(osp, kind, 2, True, p_ev, t_ev, tend, fsptol, dsinks)
        dropped += ids.size
        assert np.array_equal(out.p[first].states, osp.states)
        assert np.array_equal(out.p[first].values, p_new)
        assert np.array_equal(out.sinks[first], OA.restart_sinks(sinks_ev, t_ev, tend, fsptol))
    for t, p, s in zip(out.t, out.p, out.sinks):
        assert s.sum() <= fsptol * t / tend + 1e-13
        assert p.sum() + s.sum() <= 1.0 + 1e-12


def test_reference_parameter_free_vs_parametric(gpu):
    # test/test_solver.jl:79-100 (bursting model with a time-varying deactivation rate)
    S, props_p, th = models.telegraph_tv("sep")
    k01, k10, lam, gam = th
    free = [
        ("ti", lambda x, p: k01 * x[0]),
        ("sep", lambda t, p: nh.nmath.max(0.0, 1.0 - nh.nmath.sin(nh.nmath.pi * t / 2)), lambda x, p: k10 * x[1]),
        ("ti", lambda x, p: lam * x[1]),
        ("ti", lambda x, p: gam * x[2]),
    ]
    m_free = nh.CmeModel(S, models.to_product(free), [])
    m_par = nh.CmeModel(S, models.to_product(props_p), th)
    ts = np.arange(0.0, 121.0, 20.0)
    p0 = nh.FspVectorSparse([[1, 0, 0]], [1.0])
    alg = nh.AdaptiveFspSparse(ode_method=None, space_adapter=nh.SelectiveRStepAdapter(10, 10, True))
    o1 = nh.solve(m_free, p0, (0.0, 120.0), alg, saveat=ts, odertol=1e-4, odeatol=1e-14)
    o2 = nh.solve(m_par, p0, (0.0, 120.0), alg, saveat=ts, odertol=1e-4, odeatol=1e-14)
    for o in (o1, o2):
        for p, s in zip(o.p, o.sinks):
            assert abs(p.sum() + s.sum() - 1.0) <= 1.5e-8 * 1.0
    assert len(o1) == len(o2) == len(ts) + 1
    for p1, p2 in zip(o1.p, o2.p):
        assert np.array_equal(p1.states, p2.states)
        assert l1(p1.values, p2.values) <= 1e-14


def test_adaptive_fp32_close_to_fp64(gpu):
    S, props, th, model = _telegraph_model()
    ts = np.arange(0.0, 121.0, 20.0)
    outs = {}
    for prec in ("fp64", "fp32"):
        alg = nh.AdaptiveFspSparse(ode_method=nh.KrylovExpv(m=30, tol=1e-12, precision=prec),
                                   space_adapter=nh.RStepAdapter(3, 3, False))
        outs[prec] = nh.solve(model, nh.FspVectorSparse([[1, 0, 0]], [1.0]), (0.0, 120.0), alg, saveat=ts,
                              fsptol=1e-4)
    a, b = outs["fp64"], outs["fp32"]
    assert len(a) == len(b)
    for k in range(len(a)):
        assert l1_distance(a.p[k].states, a.p[k].values, b.p[k].states, b.p[k].values) <= 1e-6


def test_adaptive_cascade_invariants(gpu):
    S, props, th = models.cascade()
    model = nh.CmeModel(S, models.to_product(props), th)
    fsptol, tend = 1e-4, 4.0
    ts = [0.0, 1.0, 2.0, 3.0, 4.0]
    alg = nh.AdaptiveFspSparse(ode_method=nh.KrylovExpv(m=30, tol=1e-12), space_adapter=nh.RStepAdapter(5, 5, True))
    out = nh.solve(model, nh.FspVectorSparse([[0, 0, 0]], [1.0]), (0.0, tend), alg, saveat=ts, fsptol=fsptol)
    assert len(out) == len(ts) + 1 and out.stats["adaptations"] >= 1
    for t, p, s in zip(out.t, out.p, out.sinks):
        assert s.sum() <= fsptol * t / tend + 1e-13
        assert 1.0 - fsptol <= p.sum() + s.sum() <= 1.0 + 1e-10
        assert p.get_values().min() > -1e-12
