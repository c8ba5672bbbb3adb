"""Checks at the full sizes of the measured configurations, through properties that do not
depend on an oracle solve of that size:

* C3 (262,144 states, 12 parameters): probability conservation, zero total sensitivity
  (d/dtheta of the total mass), FP32 storage against FP64, and each sensitivity against
  central differences of two plain FP64 probability solves (a time-invariant parameter and
  the separable k21).
* C4 (all 512 instances of the committed table): every instance conserves probability at
  every output, and 32 instances stratified by size agree with the oracle's exact solution.
* C5 (cascade to t = 5, fsptol 1e-4): the learned adapter keeps the sink bound at every
  output and stays within 2 fsptol of the plain adapter's solution.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import models
from fspcompare import l1_distance
import numcme_hip as nh

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]


def _c3_solve(precision, theta=None):
    S, props, th = models.gene4() if theta is None else models.gene4(theta)
    states = models.gene4_states(255)
    model = nh.CmeModelWithSensitivity(nh.CmeModel(S, models.to_product(props), th))
    p0 = np.zeros(len(states))
    p0[0] = 1.0
    ic = nh.forwardsens_initial_condition(states, p0, [np.zeros(len(states)) for _ in th])
    alg = nh.AdaptiveForwardSensFspSparse(space_adapter=nh.ForwardSensRStepAdapter(0, 1, False),
                                          ode_method=nh.RK45(rtol=1e-8, atol=1e-14, precision=precision))
    ts = np.arange(0.0, 901.0, 300.0)
    out = nh.solve(model, ic, (0.0, 900.0), alg, saveat=ts, fsptol=1.0)
    assert out.stats["adaptations"] == 0
    return out, states, S, props, th


def _prob_final(S, props, th, states):
    model = nh.CmeModel(S, models.to_product(props), th)
    p0 = np.zeros(len(states))
    p0[0] = 1.0
    out = nh.solve(model, nh.FspVectorSparse(states, p0), (0.0, 900.0), nh.RK45(rtol=1e-10, atol=1e-16),
                   saveat=[900.0])
    return out.p[-1].values


def test_c3_full_size_invariants_and_central_differences(gpu):
    o64, states, S, props, th = _c3_solve("fp64")
    o32, *_ = _c3_solve("fp32")
    P = len(th)
    for k in range(len(o64)):
        assert abs(o64.p[k].values.sum() + o64.sinks[k].sum() - 1.0) <= 1e-9
        assert abs(o32.p[k].values.sum() + o32.sinks[k].sum() - 1.0) <= 1e-6
        for ip in range(P):
            assert abs(o64.S[k][ip].values.sum() + np.sum(o64.dsinks[k][ip])) <= 1e-10
        assert np.abs(o64.p[k].values - o32.p[k].values).sum() <= 1e-6
        for ip in range(P):
            a, b = o64.S[k][ip].values, o32.S[k][ip].values
            assert np.abs(a - b).max() <= 1e-5 * max(np.abs(a).max(), 1e-300), (k, ip)
    # central differences of the probability solve (k_export: time-invariant; k21: separable)
    for ip in (10, 3):
        d = 1e-4 * th[ip]
        tp, tm = th.copy(), th.copy()
        tp[ip] += d
        tm[ip] -= d
        fd = (_prob_final(S, props, tp, states) - _prob_final(S, props, tm, states)) / (2 * d)
        sens = o64.S[-1][ip].values
        assert np.abs(fd - sens).max() <= 1e-5 * np.abs(sens).max(), ip


def _c4_oracle_generator(spec):
    """The oracle's FP64 generator of one batched instance (same states, same order)."""
    from oracle.fspmatrix import FspMatrix
    from oracle.statespace import StateSpace
    th, n = np.asarray(spec["theta"], dtype=float), spec["nmax"]
    if spec["kind"] == "telegraph":
        S, props, _ = models.telegraph(th)
        states = models.telegraph_states(n)
    else:
        S, props, _ = models.toggle(th)
        states = models.box_states([n, n])
    return FspMatrix(StateSpace(S, states), models.to_oracle(props), th).to_csc(0.0)


def test_c4_full_batch_vs_oracle(gpu):
    """All 512 instances of the committed table (200..197,136 states, 14.2M in total) in one
    batched FP32 solve. Every instance conserves probability and stays non-negative at every
    output; 32 instances stratified by size (every 16th of the sorted table, the smallest
    and the largest included) are checked against the oracle's exact FP64 solution by
    uniformization (itself pinned against dense expm and expm_multiply): l1 <= 1e-6."""
    from bench import c4_instance  # the bench's instance builder (model, space, generator)
    from oracle import cpu_ref

    table = json.loads((ROOT / "tests" / "golden" / "c4_instances.json").read_text())
    sizes = np.array([t["states"] for t in table])
    assert sizes.min() == 200 and sizes.max() > 150_000 and len(table) == 512
    mats, u0s = [], []
    for spec in table:
        A, u0 = c4_instance(nh, spec, "fp32")
        mats.append(A)
        u0s.append(u0)
    B = nh.BatchFsp(mats)
    ts = np.arange(0.0, 11.0)
    B.solve(u0s, ts, fetch=False)
    print(f"batched: {B.stats}", flush=True)
    for i in range(len(table)):
        for k in range(len(ts)):
            sn = B.snapshot(i, k)
            assert abs(sn.sum() - 1.0) <= 1e-6, (i, k)
            assert sn.min() >= -1e-9, (i, k)
    order = np.argsort(sizes, kind="stable")
    sample = order[np.linspace(0, len(order) - 1, 32).round().astype(int)]
    worst = 0.0
    for i in sample:
        U = cpu_ref.uniformization(_c4_oracle_generator(table[i]), u0s[i], ts)
        err = max(float(np.abs(B.snapshot(int(i), k) - U[k]).sum()) for k in range(len(ts)))
        worst = max(worst, err)
        assert err <= 1e-6, (int(i), table[i]["kind"], table[i]["states"], err)
    print(f"C4 stratified sample: worst l1 {worst:.2e} over {len(sample)} instances", flush=True)


def test_c5_learned_full_horizon(gpu):
    import bench

    rng = np.random.default_rng(0)
    rec = nh.TrainingRecords(capacity=400_000, per_adaptation=1024, label_mass=1e-10)
    rk = lambda prec: nh.RK45(rtol=1e-6, atol=1e-10, precision=prec)
    for _ in range(16):
        th = np.asarray(bench.C5_THETA) * np.exp(rng.uniform(np.log(0.5), np.log(2.0), size=6))
        nh.solve(bench.cascade_model(nh, th), nh.FspVectorSparse([[0, 0, 0]], [1.0]), (0.0, 1.5),
                 nh.AdaptiveFspSparse(ode_method=rk("fp32"), space_adapter=nh.SelectiveRStepAdapter(2, 2, False)),
                 saveat=np.linspace(0.0, 1.5, 6), fsptol=1e-4, records=rec)
    net, _ = nh.train_adapter(*rec.arrays(), epochs=20, batch=4096, lr=2e-3, seed=0)
    model = bench.cascade_model(nh, bench.C5_THETA)
    ts = np.linspace(0.0, 5.0, 6)
    fsptol = 1e-4
    run = lambda ad, prec: nh.solve(model, nh.FspVectorSparse([[0, 0, 0]], [1.0]), (0.0, 5.0),
                                    nh.AdaptiveFspSparse(ode_method=rk(prec), space_adapter=ad), saveat=ts,
                                    fsptol=fsptol)
    learned = run(nh.LearnedAdapter(net, 2, 8, False, threshold=0.3), "fp32")
    plain = run(nh.SelectiveRStepAdapter(2, 12, False), "fp64")
    assert learned.stats["learned_adaptations"] > 0
    assert learned.stats["max_states"] < plain.stats["max_states"] / 4
    for t, p, s in zip(learned.t, learned.p, learned.sinks):
        assert s.sum() <= fsptol * t / 5.0 + 1e-12
        assert 1.0 - fsptol <= p.sum() + s.sum() <= 1.0 + 1e-6
    for a, b in zip(learned.p, plain.p):
        assert l1_distance(a.states, a.values, b.states, b.values) <= 2 * fsptol
