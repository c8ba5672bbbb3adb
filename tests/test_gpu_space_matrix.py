"""GPU parity of the device state space and FSP matrix against the oracle.

State indexing must be bit-exact (states, order, connectivity, sinks); matrix entries
match the oracle's FP64 CSC to rounding; FP32 storage keeps exact zero column sums.
"""
import numpy as np
import pytest

import models
import numcme_hip as nh
from oracle.fspmatrix import FspMatrix
from oracle.statespace import StateSpace

pytestmark = pytest.mark.gpu


def same_space(dev, ora):
    states, sc, sk = ora.arrays()
    assert np.array_equal(dev.get_states(), states)
    assert np.array_equal(dev.get_state_connectivity(), sc)
    assert np.array_equal(dev.get_sink_connectivity(), sk)


def test_reference_statespace_cases(gpu):
    # test/test_statespace.jl:42-87 plus the LIFO known-answer orders
    S = models.TELEGRAPH_S
    for lv, n in ((0, 1), (1, 3), (3, 7), (4, 9)):
        d = nh.StateSpaceSparse(S, [1, 0, 0])
        o = StateSpace(S, [[1, 0, 0]])
        d.expand(lv), o.expand(lv)
        assert d.get_state_count() == n
        same_space(d, o)
    T = models.TOGGLE_S
    for lv, n in ((0, 1), (1, 3), (3, 10)):
        d = nh.StateSpaceSparse(T, [0, 0])
        o = StateSpace(T, [[0, 0]])
        d.expand(lv), o.expand(lv)
        assert d.get_state_count() == n
        same_space(d, o)
    d = nh.StateSpaceSparse(T, [[0, 0]])
    d.expand(3)
    assert d.get_states().tolist() == [[0, 0], [1, 0], [0, 1], [1, 1], [0, 2], [2, 0], [3, 0], [2, 1], [1, 2], [0, 3]]


def test_constructor_dedup_and_negative(gpu):
    d = nh.StateSpaceSparse(models.TOGGLE_S, [[2, 1], [0, -1], [2, 1], [0, 0]])
    assert d.get_states().tolist() == [[2, 1], [0, 0]]
    same_space(d, StateSpace(models.TOGGLE_S, [[2, 1], [0, -1], [2, 1], [0, 0]]))


@pytest.mark.parametrize("seed", range(6))
def test_random_call_sequences(gpu, seed):
    rng = np.random.default_rng(seed)
    S = [models.TOGGLE_S, models.TELEGRAPH_S, models.gene4_stoich()][seed % 3]
    x0 = [[0, 0], [1, 0, 0], [1, 0, 0, 0, 0, 0]][seed % 3]
    d, o = nh.StateSpaceSparse(S, [x0]), StateSpace(S, [x0])
    for step in range(8):
        if rng.random() < 0.65 or o.count < 3:
            lv = int(rng.integers(0, 6))
            only = sorted(rng.choice(S.shape[1], size=int(rng.integers(0, S.shape[1] + 1)), replace=False) + 1)
            d.expand(lv, only), o.expand(lv, only)
        else:
            k = int(rng.integers(1, max(2, o.count // 3)))
            ids = np.sort(rng.choice(o.count, size=k, replace=False)) + 1
            d.deleteat(ids), o.deleteat(ids)
        same_space(d, o)
    q = np.vstack([o.states[:50], o.states[:5] + 1000, -np.ones((1, S.shape[0]), int)])
    assert np.array_equal(d.lookup(q), o.lookup(q))


def test_large_expansion_bit_exact(gpu):
    # ~45k states through 300 levels of toggle exploration, then a prune and re-expansion
    d, o = nh.StateSpaceSparse(models.TOGGLE_S, [0, 0]), StateSpace(models.TOGGLE_S, [[0, 0]])
    d.expand(300), o.expand(300)
    same_space(d, o)
    ids = np.arange(1, o.count + 1)[(o.states.sum(axis=1) % 7) == 3]
    d.deleteat(ids), o.deleteat(ids)
    same_space(d, o)
    d.expand(5, [1, 3]), o.expand(5, [1, 3])
    same_space(d, o)


def test_cascade_expansion_bit_exact(gpu):
    S = np.array([[1, 0, 0, -1, 0, 0], [0, 1, 0, 0, -1, 0], [0, 0, 1, 0, 0, -1]])
    d, o = nh.StateSpaceSparse(S, [0, 0, 0]), StateSpace(S, [[0, 0, 0]])
    d.expand(40), o.expand(40)
    assert o.count > 10000
    same_space(d, o)


def test_delete_all_and_empty(gpu):
    d = nh.StateSpaceSparse(models.TOGGLE_S, [0, 0])
    d.expand(2)
    d.deleteat(np.arange(1, d.get_state_count() + 1))
    assert d.get_state_count() == 0
    with pytest.raises(nh.ArgumentError):
        d.deleteat([1])
    e = nh.StateSpaceSparse(models.TOGGLE_S, [[3, 4]])
    with pytest.raises(nh.ArgumentError):
        e.deleteat([1, 1, 0])


# ---- matrix ----------------------------------------------------------------------------
def build_pair(S, props, th, states, precision="fp64"):
    d = nh.StateSpaceSparse(S, states)
    o = StateSpace(S, states)
    Ad = nh.FspMatrixSparse(d, models.to_product(props), th, precision=precision)
    Ao = FspMatrix(o, models.to_oracle(props), th)
    return d, o, Ad, Ao


def test_nnz_and_entries_telegraph_c1(gpu):
    S, props, th = models.telegraph()
    _, _, Ad, Ao = build_pair(S, props, th, models.telegraph_states(399))
    info = Ad.info()
    assert info["n_rows"] == 804 and info["nnz_reference"] == 3198 == Ao.nnz()
    diff = (Ad.to_scipy(0.0) - Ao.to_csc(0.0)).toarray()
    assert np.abs(diff).max() <= 1e-14 * abs(Ao.to_csc(0.0)).max()


def test_nnz_toggle_c2(gpu):
    S, props, th = models.toggle()
    d = nh.StateSpaceSparse(S, models.box_states([1023, 1023]))
    A = nh.FspMatrixSparse(d, models.to_product(props), th, precision="fp32")
    info = A.info()
    assert info["n_rows"] == 1048580
    assert info["nnz_reference"] == 5240832
    assert info["ell_width"] == 4
    assert info["n_sink_entries"] == 2048
    assert info["computed_col_slots"] == 4  # lattice order: every slot is row + constant offset


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_computed_columns_match_loaded_columns(gpu, precision):
    # the same box in lattice order (columns computed) and in a shuffled order (columns
    # loaded): the state rows of the two mat-vecs agree bit for bit after un-permuting
    S, props, th = models.toggle()
    states = models.box_states([57, 43])
    perm = np.random.default_rng(3).permutation(len(states))
    A1 = nh.FspMatrixSparse(nh.StateSpaceSparse(S, states), models.to_product(props), th, precision=precision)
    A2 = nh.FspMatrixSparse(nh.StateSpaceSparse(S, states[perm]), models.to_product(props), th, precision=precision)
    assert A1.info()["computed_col_slots"] == 4 and A2.info()["computed_col_slots"] == 0
    n = len(states)
    v1 = np.random.default_rng(4).random(A1.rowcount)
    v2 = np.concatenate([v1[:n][perm], v1[n:]])
    w1, w2 = nh.matvec(0.0, A1, v1), nh.matvec(0.0, A2, v2)
    assert np.array_equal(w1[:n][perm], w2[:n])
    assert np.allclose(w1[n:], w2[n:], rtol=1e-14, atol=0)


@pytest.mark.parametrize("which", ["toggle", "gene4", "telegraph_sens"])
def test_entries_and_matvec_vs_oracle(gpu, which):
    if which == "toggle":
        S, props, th = models.toggle()
        states = models.box_states([40, 35])
    elif which == "gene4":
        S, props, th = models.gene4()
        states = models.gene4_states(12)
    else:
        S, props, th = models.telegraph_sens()
        states = models.telegraph_states(80)
    _, _, Ad, Ao = build_pair(S, props, th, states)
    assert Ad.info()["nnz_reference"] == Ao.nnz()
    rng = np.random.default_rng(0)
    v = rng.random(Ao.rowcount)
    for t in (0.0, 10.0, 20.0, 133.7):
        ref = Ao.to_csc(t)
        got = Ad.to_scipy(t)
        assert abs(got - ref).max() <= 1e-14 * abs(ref).max()
        w = nh.matvec(t, Ad, v)
        wo = Ao.matvec(t, v)
        assert np.abs(w - wo).max() <= 1e-13 * np.abs(wo).max()
        out = np.ones(Ao.rowcount)
        Ad.matvecadd_(out, t, v)
        assert np.abs(out - 1.0 - wo).max() <= 1e-13 * max(1.0, np.abs(wo).max())


def test_fp32_exact_zero_column_sums(gpu):
    S, props, th = models.toggle()
    _, _, Ad, Ao = build_pair(S, props, th, models.box_states([60, 60]), precision="fp32")
    A = Ad.to_scipy(0.0)
    colsum = np.asarray(A.sum(axis=0)).ravel()
    # interior columns conserve mass exactly; every entry is within FP32 rounding of FP64
    assert np.abs(colsum).max() == 0.0
    ref = Ao.to_csc(0.0)
    rel = abs(A - ref).max() / abs(ref).max()
    assert rel < 4e-7


def test_reference_fspmat_cases(gpu):
    # test/test_fspmat.jl: zero column sums at t = 0, 1 for TI/separable/joint and
    # separable == joint at t = 0
    ws = {}
    for kind in ("ti", "sep", "joint"):
        if kind == "ti":
            S, props, th = models.telegraph((0.05, 0.1, 5.0, 1.0))
        else:
            S, props, th = models.telegraph_tv(kind)
        d = nh.StateSpaceSparse(S, [1, 0, 0])
        d.expand(2)
        A = nh.FspMatrixSparse(d, models.to_product(props), th)
        n = A.size(1)
        assert n == d.get_state_count() + d.get_sink_count() == A.size(2)
        for t in (1.0, 0.0):
            w = nh.matvec(t, A, np.ones(n))
            assert abs(w.sum()) <= 1e-14
        ws[kind] = w
    assert np.array_equal(ws["sep"], ws["joint"])


def test_sens_matrix_zero_sum_poisson(gpu):
    # test/sensmat/poisson.jl: constant-propensity birth-death on states 1..100
    S = np.array([[1, -1]])
    props = [("ti", lambda x, p: p[0]), ("ti", lambda x, p: p[1])]
    d = nh.StateSpaceSparse(S, [[i] for i in range(1, 101)])
    A = nh.FspMatrixSparse(d, models.to_product(props), [10.0, 5.0])
    w = nh.matvec(0.0, A, np.ones(A.rowcount))
    assert abs(w.sum()) < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_rows_per_thread_variants_identical(gpu, precision):
    # 803 states (not a multiple of 4): every launch shape gives the same bits
    S, props, th = models.toggle()
    states = models.box_states([72, 10])[:803]
    d = nh.StateSpaceSparse(S, states)
    o = StateSpace(S, states)
    A = nh.FspMatrixSparse(d, models.to_product(props), th, precision=precision)
    Ao = FspMatrix(o, models.to_oracle(props), th)
    v = np.random.default_rng(1).random(A.rowcount)
    outs = []
    for r in (1, 2, 4):
        A.set_rows_per_thread(r)
        outs.append(nh.matvec(0.0, A, v))
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    ref = Ao.matvec(0.0, v if precision == "fp64" else v.astype(np.float32).astype(float))
    tol = 1e-13 if precision == "fp64" else 1e-6
    assert np.abs(outs[0] - ref).max() <= tol * np.abs(ref).max()
    rows, ms = A.autotune(reps=5)
    assert rows in (1, 2, 4) and ms > 0


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_long_sink_rows_vs_oracle(gpu, precision):
    # a jump reaction (+100 in species 0) on a 128 x 128 box: 12,800 states share its sink
    # row, so the row is reduced in chunks by several workgroups and folded in order
    S = np.array([[100, -1, 0, 0], [0, 0, 1, -1]])
    props = [("ti", lambda x, p: p[0] + 0.0 * x[0]), ("ti", lambda x, p: p[1] * x[0]),
             ("ti", lambda x, p: p[2] + 0.0 * x[1]), ("ti", lambda x, p: p[3] * x[1])]
    th = np.array([0.7, 0.05, 3.0, 0.1])
    states = models.box_states([127, 127])
    _, _, Ad, Ao = build_pair(S, props, th, states, precision=precision)
    info = Ad.info()
    assert info["n_sink_entries"] > 8192
    rng = np.random.default_rng(2)
    v = rng.random(Ao.rowcount)
    vin = v if precision == "fp64" else v.astype(np.float32).astype(float)
    wo = Ao.matvec(0.0, vin)
    tol = 1e-13 if precision == "fp64" else 1e-6
    for acc in (False, True):
        if acc:
            w = np.full(Ao.rowcount, 0.5)
            Ad.matvecadd_(w, 0.0, v)
            w -= 0.5
        else:
            w = nh.matvec(0.0, Ad, v)
        assert np.abs(w - wo).max() <= tol * np.abs(wo).max()
        # the sink row of the jump: every state with x0 + 100 > 127
        assert abs(w[len(states)] - wo[len(states)]) <= tol * abs(wo[len(states)])


@pytest.mark.parametrize("order", ["lattice", "shuffled"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_twelve_slot_matvec_vs_oracle(gpu, order, precision):
    # K = 12 > 8 takes the generic kernel (runtime slot count, per-slot column kind): in
    # lattice order every slot's column is computed, in a shuffled order every one loaded
    S, props, th = models.twelve_slot()
    states = models.box_states([9, 8, 7])
    if order == "shuffled":
        states = states[np.random.default_rng(5).permutation(len(states))]
    d, o, Ad, Ao = build_pair(S, props, th, states, precision=precision)
    info = Ad.info()
    assert info["ell_width"] == 12 and info["nnz_reference"] == Ao.nnz()
    assert info["computed_col_slots"] == (12 if order == "lattice" else 0)
    v = np.random.default_rng(6).random(Ao.rowcount)
    w = nh.matvec(0.0, Ad, v)
    wo = Ao.matvec(0.0, v)
    bar = 1e-13 if precision == "fp64" else 1e-5
    assert np.abs(w - wo).max() <= bar * np.abs(wo).max()
