"""GPU parity of the batched solve: every instance of a batch of uneven telegraph and
toggle generators against the exact oracle (small instances) and the single-instance
device solve, in FP64 and FP32; graph replays against eager launches.
"""
import numpy as np
import pytest

import models
import numcme_hip as nh
from oracle.solve import solve_fixed

pytestmark = pytest.mark.gpu


def _instances(precision):
    out = []
    rng = np.random.default_rng(11)
    for k in range(6):
        th = (10 ** rng.uniform(-2, 0), 10 ** rng.uniform(-2, 0), 10 ** rng.uniform(0, 1.3), 10 ** rng.uniform(-1, 0))
        S, props, th = models.telegraph(th)
        states = models.telegraph_states(int(rng.integers(20, 120)))
        out.append((S, props, th, states))
    for k in range(5):
        th = (10 ** rng.uniform(1, 1.5), 20.0, 2.5, 1.0, 1.0)
        S, props, th = models.toggle(th)
        states = models.box_states([int(rng.integers(10, 45))] * 2)
        out.append((S, props, th, states))
    mats, u0s = [], []
    for S, props, th, states in out:
        sp = nh.StateSpaceSparse(S, states)
        mats.append(nh.FspMatrixSparse(sp, models.to_product(props), th, precision=precision))
        u0 = np.zeros(len(states) + S.shape[1])
        u0[0] = 1.0
        u0s.append(u0)
    return out, mats, u0s


@pytest.mark.parametrize("precision,bar,rows", [("fp64", 1e-8, 1), ("fp64", 1e-8, 2), ("fp64", 1e-8, 4),
                                               ("fp32", 1e-6, 4), ("fp32", 1e-6, "tuned")])
def test_batch_matches_oracle(gpu, precision, bar, rows):
    inst, mats, u0s = _instances(precision)
    B = nh.BatchFsp(mats)
    if rows == "tuned":
        r, ms = B.autotune(reps=2)
        assert r in (1, 2, 4) and ms > 0
    else:
        B.set_rows_per_thread(rows)
    ts = np.array([0.0, 0.5, 2.0, 5.0])
    snaps = B.solve(u0s, ts, rtol=1e-10, atol=1e-16)
    for (S, props, th, states), snap in zip(inst, snaps):
        p0 = np.zeros(len(states))
        p0[0] = 1.0
        ref = solve_fixed(S, models.to_oracle(props), th, states, p0, ts)
        n = len(states)
        for k in range(len(ts)):
            assert np.abs(snap[k][:n] - ref.p[k]).sum() <= bar
            assert np.abs(snap[k][n:] - ref.sinks[k]).sum() <= bar


def test_batch_graphs_equal_eager(gpu):
    inst, mats, u0s = _instances("fp32")
    B = nh.BatchFsp(mats)
    ts = np.array([0.0, 1.0, 3.0])
    a = B.solve(u0s, ts, rtol=1e-8, graphs=True)
    b = B.solve(u0s, ts, rtol=1e-8, graphs=False)
    c = B.solve(u0s, ts, rtol=1e-8, graphs=True)  # replays of the graph captured by the first call
    for x, y, z in zip(a, b, c):
        for k in range(len(ts)):
            assert np.array_equal(x[k], y[k]) and np.array_equal(x[k], z[k])
    # the one-call download of every instance equals the per-instance copies
    for k in range(len(ts)):
        every = B.snapshots(k)
        assert len(every) == len(mats)
        for i, row in enumerate(every):
            assert np.array_equal(row, B.snapshot(i, k)) and np.array_equal(row, c[i][k])
    with pytest.raises(nh.NumCmeHipError):
        B.snapshots(len(ts))


def test_batch_rejects_time_varying(gpu):
    S, props, th = models.telegraph_tv("sep")
    sp = nh.StateSpaceSparse(S, models.telegraph_states(10))
    A = nh.FspMatrixSparse(sp, models.to_product(props), th)
    with pytest.raises(nh.NumCmeHipError):
        nh.BatchFsp([A])


def test_batch_rows_per_thread_rejected(gpu):
    _, mats, _ = _instances("fp64")
    B = nh.BatchFsp(mats)
    with pytest.raises(nh.ArgumentError):
        B.set_rows_per_thread(3)


@pytest.mark.parametrize("precision,bar", [("fp64", 1e-8), ("fp32", 1e-6)])
def test_batch_twelve_slots_vs_oracle(gpu, precision, bar):
    # three-species instances with 12 distinct stoichiometries next to telegraph instances:
    # the concatenated slot width is 12, above the fixed-slot-count batched kernels
    inst = []
    for k, bounds in enumerate(([5, 4, 3], [7, 3, 2], [4, 4, 4])):
        S, props, th = models.twelve_slot()
        inst.append((S, props, th * (1.0 + 0.3 * k), models.box_states(bounds)))
    for nmax in (30, 55):
        S, props, th = models.telegraph()
        inst.append((S, props, th, models.telegraph_states(nmax)))
    mats, u0s = [], []
    for S, props, th, states in inst:
        sp = nh.StateSpaceSparse(S, states)
        mats.append(nh.FspMatrixSparse(sp, models.to_product(props), th, precision=precision))
        u0 = np.zeros(len(states) + S.shape[1])
        u0[0] = 1.0
        u0s.append(u0)
    B = nh.BatchFsp(mats)
    ts = np.array([0.0, 0.5, 2.0])
    snaps = B.solve(u0s, ts, rtol=1e-10, atol=1e-16)
    for (S, props, th, states), snap in zip(inst, snaps):
        p0 = np.zeros(len(states))
        p0[0] = 1.0
        ref = solve_fixed(S, models.to_oracle(props), th, states, p0, ts)
        n = len(states)
        for k in range(len(ts)):
            assert np.abs(snap[k][:n] - ref.p[k]).sum() <= bar
            assert np.abs(snap[k][n:] - ref.sinks[k]).sum() <= bar


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_from_models_equals_from_matrices(gpu, precision):
    # nc_batch_create_models (spaces and generators built by worker threads of the library)
    # gives the batch of the same generators: every snapshot bit-identical
    inst, mats, u0s = _instances(precision)
    cms = [nh.CmeModel(S, models.to_product(props), th) for S, props, th, _ in inst]
    B1 = nh.BatchFsp(mats)
    for workers in (1, 3):
        B2 = nh.BatchFsp.from_models(cms, [np.asarray(st) for _, _, _, st in inst], precision, workers=workers)
        assert len(B2) == len(B1) and np.array_equal(B2.rowcounts, B1.rowcounts)
        ts = np.array([0.0, 1.0, 3.0])
        s1 = B1.solve(u0s, ts, rtol=1e-9, atol=1e-15)
        s2 = B2.solve(u0s, ts, rtol=1e-9, atol=1e-15)
        for a, b in zip(s1, s2):
            for k in range(len(ts)):
                assert np.array_equal(a[k], b[k])


def test_batch_from_models_errors(gpu):
    S, props, th = models.telegraph((0.1, 0.1, 5.0, 1.0))
    cm = nh.CmeModel(S, models.to_product(props), th)
    good = np.asarray(models.telegraph_states(20))
    empty = np.zeros((0, 3), dtype=np.int32)
    with pytest.raises(nh.ArgumentError, match="instance 1"):
        nh.BatchFsp.from_models([cm, cm, cm], [good, empty, good], "fp32", workers=2)
    with pytest.raises(nh.ArgumentError):
        nh.BatchFsp.from_models([cm, cm], [good], "fp32")
    # a time-varying generator is built by the workers, then refused by the batch
    Stv, props_tv, th_tv = models.telegraph_tv("sep")
    tv = nh.CmeModel(Stv, models.to_product(props_tv), th_tv)
    with pytest.raises(nh.NumCmeHipError, match="time-invariant"):
        nh.BatchFsp.from_models([cm, tv], [good, np.asarray(models.box_states([1, 1, 20]))], "fp32", workers=2)
    # the library is usable after the failure
    B = nh.BatchFsp.from_models([cm, cm], [good, good], "fp32")
    assert len(B) == 2
