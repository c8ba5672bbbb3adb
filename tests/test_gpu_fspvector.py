"""GPU parity of the device sparse vectors (FspVectorSparse, src/fspvector/fspvector.jl)
against the oracle restatement of its loops.

The marginal ``sum(p, dims)`` must be bit-identical to the reference's loop (states in
order of first appearance, values added in state order), and ``Array(p)`` exact. Cases:
the vector of test/test_fspvec.jl, random vectors with repeated and negative coordinates,
and snapshots of fixed, adaptive and sensitivity solves taken without leaving the device.
"""
import numpy as np
import pytest

import models
import numcme_hip as nh
from oracle import fspvector as O

pytestmark = pytest.mark.gpu


def reference_vector():
    states = [[0, i] for i in range(11)] + [[1, i] for i in range(11)]
    return nh.FspVectorSparse(states, [0.0] * 11 + [1.0] * 11)


def test_reference_fspvec_cases(gpu):
    # test/test_fspvec.jl:8-14
    v = reference_vector()
    v1, v2 = v.sum([1]), v.sum([2])
    assert v.sum() == 11.0
    assert nh.nnz(v1) == 11 and nh.nnz(v2) == 2
    assert v1.states[:, 0].tolist() == list(range(11)) and v1.values.tolist() == [1.0] * 11
    assert v2.states[:, 0].tolist() == [0, 1] and v2.values.tolist() == [0.0, 11.0]
    v12 = v.sum([1, 2])
    assert v12.states.shape == (1, 0) and v12.values.tolist() == [11.0]
    # the device total of a device-resident vector
    m = v.sum([1])
    assert m.on_device() and m.sum() == 11.0 and m.on_device()


@pytest.mark.parametrize("seed", range(4))
def test_random_marginals_bit_exact(gpu, seed):
    rng = np.random.default_rng(seed)
    n, ns = 3000, 4
    st = rng.integers(-3, 9, size=(n, ns))
    st[rng.integers(0, n, 200)] = st[rng.integers(0, n, 200)]  # repeated states
    vals = rng.standard_normal(n) * 10.0 ** rng.integers(-12, 2, n)
    v = nh.FspVectorSparse(st, vals)
    for dims in ([1], [2], [4], [1, 3], [2, 4, 1], [1, 2, 3, 4], [3, 3]):
        m = v.sum(dims)
        s_ref, v_ref = O.marginal(st, vals, dims)
        assert np.array_equal(m.states, s_ref), dims
        assert np.array_equal(m.values, v_ref), dims  # the reference's order of additions


def test_dense_matches_oracle(gpu):
    rng = np.random.default_rng(9)
    st = rng.integers(0, 7, size=(400, 3))
    vals = rng.random(400)  # repeated states: the last one wins
    a = nh.FspVectorSparse(st, vals).to_array()
    ref = O.dense(st, vals)
    assert a.shape == ref.shape and np.array_equal(a, ref)
    a = nh.FspVectorSparse([[2, 0], [0, 1], [1, 1]], [0.5, 0.25, 0.25]).to_array()
    expect = np.zeros((3, 2))
    expect[2, 0], expect[0, 1], expect[1, 1] = 0.5, 0.25, 0.25
    assert np.array_equal(a, expect)
    with pytest.raises(nh.ArgumentError):
        nh.FspVectorSparse(np.zeros((0, 2), dtype=np.int64), []).to_array()


def test_fixed_solve_snapshots_stay_on_device(gpu):
    # toggle switch on a 160 x 160 box: the marginals of every snapshot are taken on the
    # device and equal the oracle's loop over the downloaded vector
    S, props, th = models.toggle()
    states = models.box_states([159, 159])
    p0 = np.zeros(len(states))
    p0[0] = 1.0
    model = nh.CmeModel(S, models.to_product(props), th)
    out = nh.solve(model, nh.FspVectorSparse(states, p0), (0.0, 2.0), nh.KrylovExpv(), saveat=[0.0, 1.0, 2.0])
    for k in range(3):
        p = out.p[k]
        assert p.on_device()
        m1, m2 = p.sum([2]), p.sum([1])
        d = p.to_array()
        assert p.on_device()  # nothing was downloaded for the marginals and the dense array
        for m, dims in ((m1, [2]), (m2, [1])):
            s_ref, v_ref = O.marginal(p.states, p.values, dims)
            assert np.array_equal(m.states, s_ref) and np.array_equal(m.values, v_ref)
        assert np.array_equal(d, O.dense(p.states, p.values))
        assert np.array_equal(p.states, states)


def test_full_size_marginals_vs_numpy(gpu):
    # 1024^2 toggle states with random values: the device marginals equal numpy's index-order
    # accumulation (np.add.at adds in index order, as the reference's loop)
    states = models.box_states([1023, 1023])
    vals = np.random.default_rng(3).random(len(states))
    v = nh.FspVectorSparse(states, vals)
    for dims, keep in (([2], 0), ([1], 1)):
        m = v.sum(dims)
        ref = np.zeros(1024)
        np.add.at(ref, states[:, keep], vals)
        assert np.array_equal(m.states[:, 0], np.arange(1024))
        assert np.array_equal(m.values, ref)
    d = v.to_array()
    assert d.shape == (1024, 1024) and np.array_equal(d[states[:, 0], states[:, 1]], vals)


def test_adaptive_and_sensitivity_snapshots(gpu):
    S, props, th = models.telegraph((0.05, 0.1, 5.0, 0.5))
    model = nh.CmeModel(S, models.to_product(props), th)
    alg = nh.AdaptiveFspSparse(ode_method=nh.RK45(rtol=1e-8, atol=1e-14),
                               space_adapter=nh.SelectiveRStepAdapter(5, 10, True))
    out = nh.solve(model, nh.FspVectorSparse([[1, 0, 0]], [1.0]), (0.0, 30.0), alg, saveat=[0.0, 10.0, 30.0],
                   fsptol=1e-6)
    for p in out.p:
        assert p.on_device()
        m = p.sum([1, 2])
        s_ref, v_ref = O.marginal(p.states, p.values, [1, 2])
        assert np.array_equal(m.states, s_ref) and np.array_equal(m.values, v_ref)
    # sensitivity snapshots: every block is a device vector over the snapshot's states
    Ss, props_s, th_s = models.telegraph_sens()
    smodel = nh.CmeModelWithSensitivity(nh.CmeModel(Ss, models.to_product(props_s), th_s))
    ic = nh.forwardsens_initial_condition([[1, 0, 0]], [1.0], [[0.0]] * len(th_s))
    salg = nh.AdaptiveForwardSensFspSparse(space_adapter=nh.ForwardSensRStepAdapter(10, 10, True),
                                           ode_method=nh.RK45(rtol=1e-8, atol=1e-14))
    sout = nh.solve(smodel, ic, (0.0, 20.0), salg, saveat=[0.0, 20.0], fsptol=1e-6)
    for k in range(len(sout.t)):
        p = sout.p[k]
        for ip in range(len(th_s)):
            sk = sout.S[k][ip]
            assert sk.on_device() and nh.nnz(sk) == nh.nnz(p)
            m = sk.sum([2])
            s_ref, v_ref = O.marginal(sk.states, sk.values, [2])
            assert np.array_equal(m.states, s_ref) and np.array_equal(m.values, v_ref)
            assert np.array_equal(sk.states, p.states)
