"""GPU tests of the learned space adapter: device features and boundary choice against the
oracle, the training records end to end, the selection path (keep-all and keep-none must
reproduce the plain selective adapter bit for bit), and a trained network keeping the sink
bound of every step.
"""
import numpy as np
import pytest

import models
from fspcompare import l1_distance
import numcme_hip as nh
from oracle import learned as ol
from oracle.statespace import StateSpace

pytestmark = pytest.mark.gpu


def _space_pair(S, init, levels):
    g = nh.StateSpaceSparse(S, init)
    g.expand(levels)
    o = StateSpace(S, init)
    o.expand(levels)
    assert np.array_equal(g.get_states(), o.states)
    return g, o


@pytest.mark.parametrize("case", ["cascade", "cascade_ds0", "telegraph_sep", "telegraph_joint"])
def test_features_match_oracle(gpu, case):
    rng = np.random.default_rng(5)
    if case.startswith("cascade"):
        S, props, th = models.cascade()
        g, o = _space_pair(S, [[0, 0, 0]], 7)
        ds = np.zeros(6) if case == "cascade_ds0" else np.array([0.3, 0.0, 0.7, 0.0, 0.0, 0.0])
        t, tend = 1.5, 6.0
    else:
        kind = "sep" if case.endswith("sep") else "joint"
        S, props, th = models.telegraph_tv(kind)
        g, o = _space_pair(S, [[1, 0, 0]], 12)
        ds = np.array([0.0, 0.0, 0.4, 0.0])
        t, tend = 0.7 if kind == "sep" else 1.3, 4.0
    model = nh.CmeModel(S, models.to_product(props), th)
    n = g.get_state_count()
    p = rng.random(n) ** 6
    p[rng.random(n) < 0.1] = 0.0
    p /= p.sum()
    ids, F = nh.boundary_features(g, model, p, ds, t, tend)
    ido, Fo = ol.features(o.states, o.sink_conn, models.to_oracle(props), th, p, ds, t, tend)
    assert ids.size > 0
    np.testing.assert_array_equal(ids, ido)
    np.testing.assert_allclose(F, Fo, rtol=0, atol=2e-6)


def _birth_death():
    S = np.array([[1, -1]])
    props = [("ti", lambda x, p: p[0] + 0.0 * x[0]), ("ti", lambda x, p: p[1] * x[0])]
    return S, props, np.array([5.0, 0.5])


def test_records_end_to_end(gpu):
    # one species: every adaptation has exactly one boundary state (the top of the space);
    # with an empty saveat the output holds the state of every event, so every record and
    # label can be recomputed by the oracle
    S, props, th = _birth_death()
    model = nh.CmeModel(S, models.to_product(props), th)
    rec = nh.TrainingRecords(capacity=1000, per_adaptation=16, label_mass=1e-7)
    alg = nh.AdaptiveFspSparse(ode_method=nh.RK45(rtol=1e-10, atol=1e-14),
                               space_adapter=nh.SelectiveRStepAdapter(2, 2, False))
    tend = 6.0
    out = nh.solve(model, nh.FspVectorSparse([[0]], [1.0]), (0.0, tend), alg, saveat=(), fsptol=1e-5, records=rec)
    na = out.stats["adaptations"]
    assert na >= 3
    X, Y = rec.arrays()
    assert len(X) == na
    sid = np.asarray(out.space_ids)
    Xo, Yo = [], []
    for j in range(na):
        last = np.nonzero(sid == j)[0][-1]        # the event that ends space j
        sl = out[last]
        states, p = np.asarray(sl.p.states), np.asarray(sl.p.values)
        sink_conn = np.zeros((len(states), 2), dtype=int)
        top = int(states[:, 0].argmax())
        sink_conn[top, 0] = 1
        ds = np.array([th[0] * p[top], 0.0])
        ids, F = ol.features(states, sink_conn, models.to_oracle(props), th, p, ds, sl.t, tend)
        Xo.append(F)
        # labelled by the state of the next event (or the end of the solve)
        nxt = np.nonzero(sid == j + 1)[0][-1]
        s2 = out[nxt]
        sp2 = StateSpace(S, np.asarray(s2.p.states))
        Yo.append(ol.labels(sp2.lookup, np.asarray(s2.p.values), S, states[ids], sink_conn[ids] != 0, 1e-7))
    np.testing.assert_allclose(X, np.concatenate(Xo), rtol=0, atol=2e-6)
    np.testing.assert_array_equal(Y, np.concatenate(Yo))


def _cascade_solve(adapter, tend=2.0, fsptol=1e-4, records=None):
    S, props, th = models.cascade()
    model = nh.CmeModel(S, models.to_product(props), th)
    alg = nh.AdaptiveFspSparse(ode_method=nh.KrylovExpv(m=30, tol=1e-10), space_adapter=adapter)
    return nh.solve(model, nh.FspVectorSparse([[0, 0, 0]], [1.0]), (0.0, tend), alg,
                    saveat=[0.0, 0.5, 1.0, 1.5, 2.0], fsptol=fsptol, records=records)


def _same(a, b):
    assert len(a) == len(b)
    for k in range(len(a)):
        assert a.t[k] == b.t[k]
        assert np.array_equal(a.p[k].states, b.p[k].states)
        assert np.array_equal(a.p[k].values, b.p[k].values)
        assert np.array_equal(a.sinks[k], b.sinks[k])


@pytest.mark.parametrize("threshold", [0.0, 1.0])
def test_keep_all_and_keep_none_equal_plain(gpu, threshold):
    plain = _cascade_solve(nh.SelectiveRStepAdapter(2, 2, True))
    net = nh.AdapterMlp(seed=1)
    learned = _cascade_solve(nh.LearnedAdapter(net, 2, 2, True, threshold=threshold))
    _same(plain, learned)
    st = learned.stats
    assert st["adaptations"] == plain.stats["adaptations"] > 0
    if threshold == 0.0:
        # every boundary state kept: the network selected every adaptation, except those
        # that followed a learned one too closely
        assert st["learned_adaptations"] > 0
    else:
        # nothing kept: every adaptation falls back to the plain adapter
        assert st["learned_adaptations"] == 0 and st["fallback_adaptations"] == st["adaptations"]


def test_trained_adapter_keeps_the_sink_bound(gpu):
    rec = nh.TrainingRecords(capacity=200000, per_adaptation=4096, label_mass=1e-8)
    plain = _cascade_solve(nh.SelectiveRStepAdapter(2, 2, False), records=rec)
    X, Y = rec.arrays()
    assert len(X) > 1000 and 0 < Y.mean() < 1
    _, l0 = nh.AdapterMlp(0).grad(X, Y)
    net, loss = nh.train_adapter(X, Y, epochs=30, batch=1024, lr=3e-3, seed=0)
    assert loss < 0.8 * l0
    fsptol, tend = 1e-4, 2.0
    learned = _cascade_solve(nh.LearnedAdapter(net, 2, 2, False, threshold=0.5), fsptol=fsptol)
    st = learned.stats
    assert st["learned_adaptations"] > 0
    for t, p, s in zip(learned.t, learned.p, learned.sinks):
        assert s.sum() <= fsptol * t / tend + 1e-13
        assert 1.0 - fsptol <= p.sum() + s.sum() <= 1.0 + 1e-9
    # both solutions are within fsptol of the exact one in l1 (plus integration error)
    d = l1_distance(plain.p[-1].states, plain.p[-1].values, learned.p[-1].states, learned.p[-1].values)
    assert d <= 2 * fsptol + 1e-8


def test_learned_adapter_argument_errors(gpu):
    S, props, th = models.cascade()
    sp = nh.StateSpaceSparse(S, [[0, 0, 0]])
    net = nh.AdapterMlp()
    with pytest.raises(nh.ArgumentError):
        nh.adapt_(sp, nh.LearnedAdapter(net, 1, 1), [1.0], np.zeros(6), 0.5, 1.0, 1e-4, dsinks=np.ones(6))
    model = nh.CmeModel(S, models.to_product(props), th)
    with pytest.raises(nh.ArgumentError):
        nh.boundary_features(sp, model, [1.0, 0.0], np.ones(6), 0.0, 1.0)
