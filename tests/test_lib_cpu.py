"""Host-side checks of the library that need no GPU.

* the shared library loads and exports every function ``include/numcme_hip.h`` declares;
* without a device, device entry points fail loudly (no CPU fallback);
* the host Hessenberg exponential (``nc_expm_dense``) matches scipy's expm;
* propensity lowering: traced programs, evaluated by an independent postfix interpreter
  written here, reproduce the callables; closed-form parameter derivatives match the
  analytic gradients of test/test_autodiff.jl:36-39 and central finite differences;
  the dependency pattern matches test_autodiff.jl:56.
"""
import ctypes
import math
import re
from pathlib import Path

import numpy as np
import pytest
import scipy.linalg as sla

import models
import numcme_hip as nh
from numcme_hip import _lib as L
from numcme_hip import expr as E

ROOT = Path(__file__).resolve().parents[1]


def header_functions():
    text = (ROOT / "include" / "numcme_hip.h").read_text()
    return sorted(set(re.findall(r"^int\s+(nc_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_header_function():
    names = header_functions()
    assert len(names) >= 30
    lib = ctypes.CDLL(str(L.LIB_PATH))
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(L.SIGNATURES), "ctypes signature table out of sync with the header"


def test_no_cpu_fallback_without_device():
    if _device_present():
        pytest.skip("a HIP device is present")
    with pytest.raises(nh.NumCmeHipError) as ei:
        nh.StateSpaceSparse([[1, -1]], [[0]])
    assert ei.value.code == L.NC_ERR_NO_DEVICE


def _device_present():
    try:
        nh.device_name()
        return True
    except nh.NumCmeHipError:
        return False


@pytest.mark.parametrize("n,scale", [(1, 1.0), (5, 0.1), (12, 3.0), (32, 40.0)])
def test_expm_dense_matches_scipy(n, scale):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) * scale / math.sqrt(n)
    ref = sla.expm(A)
    got = nh.expm_dense(A)
    assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()) * max(1.0, scale)


@pytest.mark.parametrize("scale", [1.0, 900.0, 1e5])
def test_expm_dense_markov_generator(scale):
    # large-norm generator (zero column sums): exp stays a stochastic matrix
    rng = np.random.default_rng(5)
    Q = rng.random((32, 32))
    np.fill_diagonal(Q, 0.0)
    Q -= np.diag(Q.sum(axis=0))
    ref = sla.expm(scale * Q)
    got = nh.expm_dense(scale * Q)
    assert np.abs(got - ref).max() <= 1e-11
    assert np.allclose(got.sum(axis=0), 1.0, atol=1e-11)


def test_expm_dense_hessenberg_generator():
    # an upper Hessenberg Krylov-like matrix with a large norm (many squarings)
    rng = np.random.default_rng(7)
    H = np.triu(rng.standard_normal((32, 32)), -1) * 5.0
    for tau in (1e-3, 0.1, 2.0):
        ref = sla.expm(tau * H)
        got = nh.expm_dense(tau * H)
        assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()


# ---- propensity lowering ----------------------------------------------------------------
def run_postfix(prog, t, x, p):
    """Independent evaluator of the C ABI postfix programs (include/numcme_hip.h opcodes)."""
    st = []
    for op, arg, c in prog:
        if op == 0: st.append(c)
        elif op == 1: st.append(p[arg])
        elif op == 2: st.append(float(x[arg]))
        elif op == 3: st.append(t)
        elif op == 9: st.append(-st.pop())
        elif op in (10, 11, 12, 13, 14, 15):
            f = {10: math.exp, 11: math.log, 12: math.sin, 13: math.cos, 14: math.sqrt, 15: abs}[op]
            st.append(f(st.pop()))
        elif op == 22:
            b, a, cc = st.pop(), st.pop(), st.pop()
            st.append(a if cc != 0 else b)
        else:
            b, a = st.pop(), st.pop()
            st.append({4: lambda: a + b, 5: lambda: a - b, 6: lambda: a * b, 7: lambda: a / b,
                       8: lambda: a ** b, 16: lambda: a if b < a else b, 17: lambda: b if b < a else a,
                       18: lambda: float(a < b), 19: lambda: float(a <= b), 20: lambda: float(a > b),
                       21: lambda: float(a >= b)}[op]())
    assert len(st) == 1
    return st[0]


def all_model_props():
    for fn in (models.telegraph, models.toggle, models.gene4, models.telegraph_sens):
        S, props, th = fn()
        yield S, props, th


def test_lowering_reproduces_callables():
    rng = np.random.default_rng(3)
    for S, props, th in all_model_props():
        ns = S.shape[0]
        t, xs, ps = E.symbols(ns, len(th))
        for pr in props:
            X = rng.integers(0, 40, size=(25, ns))
            cols = [X[:, s] for s in range(ns)]
            for tt in (0.0, 7.3, 61.0):
                if pr[0] == "ti":
                    ref = np.broadcast_to(pr[1](cols, th), (25,))
                    prog = E.compile_expr(E.trace(pr[1], xs, ps))
                elif pr[0] == "sep":
                    ref = np.broadcast_to(pr[1](tt, th) * np.asarray(pr[2](cols, th)), (25,))
                    prog_t = E.compile_expr(E.trace(pr[1], t, ps))
                    prog_x = E.compile_expr(E.trace(pr[2], xs, ps))
                    got = [run_postfix(prog_t, tt, X[i], th) * run_postfix(prog_x, tt, X[i], th) for i in range(25)]
                    assert np.allclose(got, ref, rtol=1e-14, atol=0)
                    continue
                else:
                    ref = np.broadcast_to(pr[1](tt, cols, th), (25,))
                    prog = E.compile_expr(E.trace(pr[1], t, xs, ps))
                got = [run_postfix(prog, tt, X[i], th) for i in range(25)]
                assert np.allclose(got, ref, rtol=1e-14, atol=0)


def test_gradients_match_reference_analytic():
    # test_autodiff.jl:7-53 model and closed-form gradients
    S, props, th = models.telegraph_sens()
    m = nh.CmeModelWithSensitivity(nh.CmeModel(S, models.to_product(props), th))
    rng = np.random.default_rng(11)
    X = rng.integers(0, 100, size=(50, 3))
    cols = [X[:, s] for s in range(3)]
    g0 = m.propensity_gradient(0, cols)
    assert np.allclose(g0[0], X[:, 0]) and all(np.all(g == 0) for g in g0[1:])
    g1 = m.propensity_gradient(1, cols)
    assert np.allclose(g1[1], X[:, 1])
    g2 = m.propensity_gradient(2, cols)
    assert np.allclose(g2[2], X[:, 1])
    for t in np.arange(0.0, 41.0):
        g3 = m.propensity_gradient(3, cols, t=t)
        L_, gam = th[4], th[3]
        s = 1.0 - math.sin(math.pi * t / L_)
        if s > 0:
            ref4 = s * X[:, 2]
            ref5 = gam * X[:, 2] * math.cos(math.pi * t / L_) * math.pi * t / L_ ** 2
        else:
            ref4 = ref5 = 0 * X[:, 2]
        assert np.allclose(g3[3], ref4, rtol=1e-8, atol=10 * np.finfo(float).eps)
        assert np.allclose(g3[4], ref5, rtol=1e-8, atol=10 * np.finfo(float).eps)


def test_sparsity_pattern_matches_reference():
    # test_autodiff.jl:56: sparse([1,2,3,4,4],[1,2,3,4,5],true)
    S, props, th = models.telegraph_sens()
    m = nh.CmeModelWithSensitivity(nh.CmeModel(S, models.to_product(props), th))
    ref = np.zeros((4, 5), dtype=bool)
    for r, k in zip([0, 1, 2, 3, 3], [0, 1, 2, 3, 4]):
        ref[r, k] = True
    assert np.array_equal(m.get_gradient_sparsity_patterns(), ref)


def test_gradients_vs_finite_differences():
    S, props, th = models.toggle()
    m = nh.CmeModelWithSensitivity(nh.CmeModel(S, models.to_product(props), th))
    X = np.array([[0, 0], [3, 17], [40, 9], [100, 100]])
    cols = [X[:, 0], X[:, 1]]
    for r, pr in enumerate(props):
        g = m.propensity_gradient(r, cols)
        for k in range(len(th)):
            h = 1e-6 * max(1.0, abs(th[k]))
            tp, tm = th.copy(), th.copy()
            tp[k] += h
            tm[k] -= h
            fd = (np.asarray(pr[1](cols, tp), float) - np.asarray(pr[1](cols, tm), float)) / (2 * h)
            assert np.allclose(g[k], fd, rtol=1e-6, atol=1e-8), (r, k)


def test_control_flow_in_propensity_is_rejected():
    with pytest.raises(TypeError):
        E.trace(lambda x, p: p[0] if x[0] > 1 else 0.0, *E.symbols(2, 1)[1:])


def test_propensity_arity_dispatch():
    assert isinstance(nh.propensity(lambda x, p: x[0]), nh.StandardTimeInvariantPropensity)
    assert isinstance(nh.propensity(lambda t, x, p: t * x[0]), nh.JointTimeVaryingPropensity)
    sep = nh.propensity(lambda x, p: x[0], lambda t, p: t)
    assert isinstance(sep, nh.SeparableTimeVaryingPropensity) and nh.istimeseparable(sep)
    with pytest.raises(nh.ArgumentError):
        nh.propensity(lambda a: a)


def _joint_laws():
    """(law, expected lowering) pairs for the joint -> separable factoring."""
    sig = lambda t: (1.0 - E.exp(-6e-3 * t)) * E.exp(-9e-3 * t)
    return [
        (lambda t, x, p: E.max(0.0, 1.0 - E.sin(E.pi * t / 2)) * p[1] * x[1], "sep"),  # test_fspmat.jl:113
        (lambda t, x, p: E.max(0.0, p[3] - 0.12 * sig(t)) * x[2], "sep"),  # the C3 k21(t) law
        (lambda t, x, p: p[0] * x[0] / (1.0 + t), "sep"),
        (lambda t, x, p: -(p[0] * (1.0 - t)) * (x[0] * x[1]) / (p[1] + x[0]), "sep"),
        (lambda t, x, p: p[0] * x[0] + 0.0 * t, "joint"),  # a sum: not a product
        (lambda t, x, p: p[0] * x[0] / (1.0 + t * x[0]), "joint"),
        (lambda t, x, p: p[1] * x[0] * x[1], "ti"),  # reads no time
    ]


def test_joint_factoring_kinds_and_values():
    rng = np.random.default_rng(21)
    t, xs, ps = E.symbols(3, 4)
    th = np.array([0.7, 1.3, 0.4, 2e-2])
    X = rng.integers(0, 50, size=(40, 3))
    for law, kind in _joint_laws():
        e = E.trace(law, t, xs, ps)
        split = E.factor_time(e)
        if kind == "joint":
            assert split is None
            continue
        assert split[0] == kind
        for tt in (0.0, 0.37, 1.0, 5.5, 120.0):
            for i in range(len(X)):
                ref = run_postfix(E.compile_expr(e), tt, X[i], th)
                if kind == "ti":
                    got = run_postfix(E.compile_expr(split[1]), tt, X[i], th)
                else:
                    tf, sf = split[1], split[2]
                    assert not E.uses(tf, "SPECIES") and not E.uses(sf, "TIME")
                    got = run_postfix(E.compile_expr(tf), tt, X[i], th) * run_postfix(E.compile_expr(sf), tt, X[i], th)
                # the factored product only reorders a product: a few ulps apart
                assert abs(got - ref) <= 4 * np.finfo(float).eps * abs(ref), (kind, tt, X[i])


def test_model_lowers_factorable_joint_laws():
    S = np.array([[1, -1, 0, 0, 0, 0, 0], [0, 0, 1, -1, 0, 0, 0], [0, 0, 0, 0, 1, -1, 1]])
    props = [nh.propensity(law) for law, _ in _joint_laws()]
    m = nh.CmeModel(S, props, [0.7, 1.3, 0.4, 2e-2])
    kinds = {"sep": L.NC_PROP_SEPARABLE, "joint": L.NC_PROP_JOINT, "ti": L.NC_PROP_TIME_INVARIANT}
    assert m.programs.kinds == [kinds[k] for _, k in _joint_laws()]
    assert m.programs.factored == [None if k == "joint" else L.NC_PROP_JOINT for _, k in _joint_laws()]
    kept = nh.CmeModel(S, props, [0.7, 1.3, 0.4, 2e-2], factor_joint=False)
    assert kept.programs.kinds == [L.NC_PROP_JOINT] * len(props)


def test_factored_joint_gradients_vs_finite_differences():
    # the separable gradient rule sf * dtf + tf * dsf (propensitygrad.jl:56-67) applied to a
    # factored joint law must be the derivative of the joint law itself
    law = lambda t, x, p: E.max(0.0, p[3] - 0.12 * (1.0 - E.exp(-6e-3 * t)) * E.exp(-9e-3 * t)) * p[1] * x[2]
    S = np.array([[0], [0], [-1]])
    th = np.array([0.7, 1.3, 0.4, 2e-2])
    m = nh.CmeModelWithSensitivity(nh.CmeModel(S, [nh.propensity(law)], th))
    assert m.cmemodel.programs.kinds == [L.NC_PROP_SEPARABLE]
    X = np.array([[0, 0, 0], [1, 2, 3], [5, 0, 40]])
    cols = [X[:, s] for s in range(3)]
    for tt in (0.0, 30.0, 400.0):
        g = m.propensity_gradient(0, cols, t=tt)
        for k in range(len(th)):
            h = 1e-6 * max(1.0, abs(th[k]))
            tp, tm = th.copy(), th.copy()
            tp[k] += h
            tm[k] -= h
            fd = (np.asarray(law(tt, cols, tp), float) - np.asarray(law(tt, cols, tm), float)) / (2 * h)
            assert np.allclose(g[k], fd, rtol=1e-6, atol=1e-10), (tt, k)


def test_with_parameters_shares_programs():
    # a parameter sweep re-uses the traced programs; only the values change
    S, props, th = models.toggle()
    m = nh.CmeModel(S, models.to_product(props), th)
    th2 = np.asarray(th, float) * 1.5
    m2 = m.with_parameters(th2)
    assert m2.programs is m.programs and np.array_equal(m2.parameters, th2)
    assert np.array_equal(m.parameters, np.asarray(th, float))
    X = np.array([[0, 0], [3, 7], [40, 2]])
    for r, (_, f) in enumerate(props):
        for x in X:
            ref = f(x, th2)
            got = run_postfix(E.compile_expr(m2.programs.xexpr[r]), 0.0, x, th2)
            assert abs(got - ref) <= 4 * np.finfo(float).eps * abs(ref)
    with pytest.raises(nh.ArgumentError):
        m.with_parameters(th2[:-1])
