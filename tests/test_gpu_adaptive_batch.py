"""GPU tests of the batched adaptive solve (``nc_solve_adaptive_batch``): many adaptive
solves run by worker threads of the library at once.

* Every instance of a batch against the oracle's exact adaptive solve (the reference's
  adaptive driver, fspsolve.jl:105-197, with the R-step adapter of rstepadapters.jl:23-52):
  the same state lists, probabilities within l1 1e-8 and sinks within 1e-10, as the single
  solve is held to.
* Each instance of a batch bit-identical to the single solve of that instance (plain
  selective, learned, FP64 and FP32, Krylov and DP5(4)), whatever the number of workers
  and the order in which the workers pick the instances.
* Errors name the instance that failed and return no output.
"""
import numpy as np
import pytest

import models
import numcme_hip as nh
from oracle import adaptive as OA

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert len(a) == len(b)
    assert a.space_ids == b.space_ids
    for k in range(len(a)):
        assert a.t[k] == b.t[k]
        assert np.array_equal(a.p[k].states, b.p[k].states)
        assert np.array_equal(a.p[k].values, b.p[k].values)
        assert np.array_equal(a.sinks[k], b.sinks[k])
    for f in ("steps", "rejected", "matvecs", "adaptations", "learned_adaptations", "fallback_adaptations",
              "max_states"):
        assert a.stats[f] == b.stats[f], f


TELEGRAPH_THETAS = [(0.05, 0.1, 5.0, 1.0), (0.2, 0.1, 2.0, 0.5), (0.05, 0.3, 10.0, 1.0), (0.5, 0.5, 3.0, 0.25),
                    (0.01, 0.02, 8.0, 2.0), (0.1, 0.1, 1.0, 0.1)]


def test_batch_matches_oracle(gpu):
    ts = np.arange(0.0, 61.0, 20.0)
    fsptol = 1e-4
    built = [models.telegraph(th) for th in TELEGRAPH_THETAS]
    cmodels = [nh.CmeModel(S, models.to_product(props), th) for S, props, th in built]
    alg = nh.AdaptiveFspSparse(ode_method=nh.KrylovExpv(m=30, tol=1e-13), space_adapter=nh.RStepAdapter(3, 3, False))
    outs = nh.solve_adaptive_batch(cmodels, nh.FspVectorSparse([[1, 0, 0]], [1.0]), (0.0, 60.0), alg, saveat=ts,
                                   fsptol=fsptol, workers=3)
    assert len(outs) == len(cmodels)
    for i, ((S, props, th), out) in enumerate(zip(built, outs)):
        ref = OA.solve_adaptive(S, models.to_oracle(props), th, [[1, 0, 0]], [1.0], (0.0, 60.0), OA.RSTEP, 3, 3, False,
                                ts, fsptol)
        print(f"instance {i}: {out.stats['adaptations']} adaptations, {out.stats['max_states']} states", flush=True)
        assert len(out) == len(ref) == len(ts) + 1
        assert out.stats["adaptations"] == len(ref.events)
        for k in range(len(out)):
            assert out.t[k] == ref.t[k]
            assert np.array_equal(out.p[k].states, ref.states[k])
            assert np.abs(out.p[k].values - ref.p[k]).sum() <= 1e-8
            assert np.abs(out.sinks[k] - ref.sinks[k]).sum() <= 1e-10


def _cascades(n, seed):
    rng = np.random.default_rng(seed)
    S, props, th = models.cascade()
    return [nh.CmeModel(S, models.to_product(props), th * np.exp(rng.uniform(np.log(0.5), np.log(2.0), size=6)))
            for _ in range(n)]


@pytest.mark.parametrize("method,precision", [("rk45", "fp64"), ("rk45", "fp32"), ("krylov", "fp64")])
def test_batch_equals_single_solves(gpu, method, precision):
    cms = _cascades(7, seed=4)
    ode = (nh.RK45(rtol=1e-6, atol=1e-10, precision=precision) if method == "rk45"
           else nh.KrylovExpv(m=30, tol=1e-10, precision=precision))
    alg = nh.AdaptiveFspSparse(ode_method=ode, space_adapter=nh.SelectiveRStepAdapter(2, 3, True))
    p0 = nh.FspVectorSparse([[0, 0, 0]], [1.0])
    kw = dict(saveat=[0.0, 0.5, 1.0, 1.5], fsptol=1e-4)
    single = [nh.solve(m, p0, (0.0, 1.5), alg, **kw) for m in cms]
    for workers in (1, 4):
        batch = nh.solve_adaptive_batch(cms, p0, (0.0, 1.5), alg, workers=workers, **kw)
        for a, b in zip(single, batch):
            _same(a, b)
    # the instances in another order: each result follows its own instance
    perm = [3, 6, 0, 5, 1, 4, 2]
    batch = nh.solve_adaptive_batch([cms[i] for i in perm], p0, (0.0, 1.5), alg, workers=3, **kw)
    for j, i in enumerate(perm):
        _same(single[i], batch[j])


def test_learned_batch_equals_single_solves(gpu):
    S, props, th = models.cascade()
    rec = nh.TrainingRecords(capacity=200000, per_adaptation=4096, label_mass=1e-8)
    plain_alg = nh.AdaptiveFspSparse(ode_method=nh.KrylovExpv(m=30, tol=1e-10),
                                     space_adapter=nh.SelectiveRStepAdapter(2, 2, False))
    p0 = nh.FspVectorSparse([[0, 0, 0]], [1.0])
    nh.solve(nh.CmeModel(S, models.to_product(props), th), p0, (0.0, 2.0), plain_alg, saveat=[0.0, 1.0, 2.0],
             fsptol=1e-4, records=rec)
    X, Y = rec.arrays()
    net, _ = nh.train_adapter(X, Y, epochs=20, batch=1024, lr=3e-3, seed=0)
    cms = _cascades(6, seed=11)
    alg = nh.AdaptiveFspSparse(ode_method=nh.RK45(rtol=1e-6, atol=1e-10),
                               space_adapter=nh.LearnedAdapter(net, 2, 4, False, threshold=0.3))
    kw = dict(saveat=[0.0, 1.0, 2.0], fsptol=1e-4)
    single = [nh.solve(m, p0, (0.0, 2.0), alg, **kw) for m in cms]
    batch = nh.solve_adaptive_batch(cms, p0, (0.0, 2.0), alg, workers=4, **kw)
    assert sum(o.stats["learned_adaptations"] for o in single) > 0
    for a, b in zip(single, batch):
        _same(a, b)
        for t, p, s in zip(b.t, b.p, b.sinks):
            assert s.sum() <= 1e-4 * t / 2.0 + 1e-13


def test_batch_errors(gpu):
    cms = _cascades(3, seed=2)
    alg = nh.AdaptiveFspSparse(ode_method=nh.RK45(rtol=1e-6, atol=1e-10),
                               space_adapter=nh.SelectiveRStepAdapter(2, 2, False))
    good = nh.FspVectorSparse([[0, 0, 0]], [1.0])
    bad = nh.FspVectorSparse([[0, 0, 0], [0, 0, 0]], [0.5, 0.5])  # repeated state
    with pytest.raises(nh.ArgumentError, match="instance 1"):
        nh.solve_adaptive_batch(cms, [good, bad, good], (0.0, 1.0), alg, saveat=[1.0], fsptol=1e-4, workers=2)
    with pytest.raises(nh.ArgumentError):
        nh.solve_adaptive_batch(cms, [good, good], (0.0, 1.0), alg, saveat=[1.0], fsptol=1e-4)
    assert nh.solve_adaptive_batch([], good, (0.0, 1.0), alg) == []
    # the library is usable after a failed batch
    out = nh.solve_adaptive_batch(cms, good, (0.0, 1.0), alg, saveat=[1.0], fsptol=1e-4, workers=2)
    assert len(out) == 3 and all(abs(o.p[-1].sum() + o.sinks[-1].sum() - 1.0) < 1e-6 for o in out)
