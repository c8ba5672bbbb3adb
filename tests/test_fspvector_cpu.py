"""FspVectorSparse: host-side construction and argument checks of the product, and the
oracle restatement of src/fspvector/fspvector.jl:57-129 pinned by the cases of
test/test_fspvec.jl, hand-computed marginals and dense arrays, and an independent numpy
group-by. No device call is made (the device marginal and dense conversion are compared
with this oracle in test_gpu_fspvector.py)."""
import numpy as np
import pytest

import numcme_hip as nh
from numcme_hip._lib import ArgumentError
from oracle import fspvector as O


def _reference_vector():
    # test_fspvec.jl:5-7: states (0, 0..10) with value 0 and (1, 0..10) with value 1
    states = [[0, i] for i in range(11)] + [[1, i] for i in range(11)]
    vals = [0.0] * 11 + [1.0] * 11
    return np.asarray(states), np.asarray(vals)


def test_reference_fspvec_cases_oracle():
    st, vals = _reference_vector()
    v1, v2 = O.marginal(st, vals, [1]), O.marginal(st, vals, [2])
    assert O.total(vals) == 11.0  # test_fspvec.jl:12
    assert len(v1[1]) == 11 and len(v2[1]) == 2  # :13-14


def test_product_constructor_errors():
    # test_fspvec.jl:16-17: unequal state and value lists
    with pytest.raises(ArgumentError):
        nh.FspVectorSparse([[0, 1]], [1.0, 2.0])
    with pytest.raises(ArgumentError):
        nh.FspVectorSparse([[0, 1], [1, 2]], [1.0])
    st, vals = _reference_vector()
    v = nh.FspVectorSparse(st, vals)
    assert v.sum() == 11.0 and nh.nnz(v) == 22
    assert v.state2idx[(1, 0)] == 12  # 1-based, as the reference's state dictionary


def test_product_dimension_errors_on_host():
    # checked before any device call (fspvector.jl:68)
    st, vals = _reference_vector()
    v = nh.FspVectorSparse(st, vals)
    for bad in ([0], [3], [1, 3], []):
        with pytest.raises(ArgumentError):
            v.sum(bad)
    with pytest.raises(ArgumentError):
        nh.FspVectorSparse(np.zeros((0, 2), dtype=np.int64), []).to_array()


def test_oracle_marginals_by_hand():
    st, vals = _reference_vector()
    # summing out species 1 leaves species 2's counts 0..10, each with 0 + 1
    s1, v1 = O.marginal(st, vals, [1])
    assert s1[:, 0].tolist() == list(range(11)) and v1.tolist() == [1.0] * 11
    # summing out species 2 leaves species 1 in order of first appearance: 0, then 1
    s2, v2 = O.marginal(st, vals, [2])
    assert s2[:, 0].tolist() == [0, 1] and v2.tolist() == [0.0, 11.0]
    # summing out both leaves one empty-coordinate state with the total
    s12, v12 = O.marginal(st, vals, [1, 2])
    assert s12.shape == (1, 0) and v12.tolist() == [11.0]
    # a repeated dimension counts once
    assert O.marginal(st, vals, [2, 2])[1].tolist() == [0.0, 11.0]
    with pytest.raises(O.ArgumentError):
        O.marginal(st, vals, [3])


def test_oracle_first_appearance_order():
    s, v = O.marginal([[3, 0], [1, 5], [3, 2], [0, 1], [1, 1]], [0.1, 0.2, 0.3, 0.15, 0.25], [2])
    assert s[:, 0].tolist() == [3, 1, 0]
    # values add in state order: (0.1 + 0.3), (0.2 + 0.25), 0.15
    assert v.tolist() == [0.1 + 0.3, 0.2 + 0.25, 0.15]


def test_oracle_marginal_vs_numpy_groupby():
    rng = np.random.default_rng(4)
    st = rng.integers(0, 6, size=(300, 3))
    vals = rng.random(300)
    for dims in ([1], [2], [3], [1, 3], [2, 3], [1, 2, 3]):
        s, v = O.marginal(st, vals, dims)
        keep = [k for k in range(3) if k + 1 not in dims]
        red = st[:, keep]
        if not keep:
            acc = vals[0]
            for x in vals[1:]:
                acc += x  # left fold in state order
            assert s.shape == (1, 0) and v.tolist() == [acc]
            continue
        _, first, inv = np.unique(red, axis=0, return_index=True, return_inverse=True)
        order = np.argsort(first, kind="stable")
        rank = np.empty_like(order)
        rank[order] = np.arange(order.size)
        # np.add.at adds in index order, as the reference's loop
        acc = np.full(order.size, np.nan)
        for i in range(len(vals)):
            g = rank[inv.ravel()[i]]
            acc[g] = vals[i] if np.isnan(acc[g]) else acc[g] + vals[i]
        assert np.array_equal(s, red[np.sort(first)])
        assert np.array_equal(v, acc)


def test_oracle_dense_array():
    a = O.dense([[2, 0], [0, 1], [1, 1]], [0.5, 0.25, 0.25])
    expect = np.zeros((3, 2))
    expect[2, 0], expect[0, 1], expect[1, 1] = 0.5, 0.25, 0.25
    assert np.array_equal(a, expect)
    # a repeated state keeps the last value (assignment in state order)
    assert O.dense([[1], [1]], [0.3, 0.7]).tolist() == [0.0, 0.7]
    with pytest.raises(O.ArgumentError):
        O.dense(np.zeros((0, 2), dtype=np.int64), [])


def test_output_indexing():
    """FspOutputSparse indexing (fspoutput.jl:44-58, 0-based here): slices, index lists,
    the last entry, and the reference's error past the end."""
    out = nh.FspOutputSparse()
    for k in range(3):
        out.t.append(float(k))
        out.p.append(nh.FspVectorSparse([[k]], [1.0]))
        out.sinks.append(np.zeros(2))
    assert len(out) == 3
    s = out[1]
    assert s.t == 1.0 and s.p.states.tolist() == [[1]] and s.sinks.tolist() == [0.0, 0.0]
    assert [x.t for x in out[[2, 0]]] == [2.0, 0.0]
    assert out[-1].t == 2.0
    with pytest.raises(ArgumentError, match="Requested index exceeds array limit."):
        out[3]


def test_saveat_step_range():
    """A numeric saveat is Julia's range tspan[1]:saveat:tspan[2], whose float elements come
    from exact rationals: 0:0.1:1 holds 0.3 and 0.7, not 3 * 0.1 and 7 * 0.1."""
    from numcme_hip.adaptive import _saveat_points

    r = _saveat_points(0.1, (0.0, 1.0))
    assert r.tolist() == [0.0, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1.0]
    assert 3 * 0.1 != 0.3  # the naive product is off by one ulp
    assert _saveat_points(20, (0.0, 120.0)).tolist() == [0.0, 20.0, 40.0, 60.0, 80.0, 100.0, 120.0]
    assert _saveat_points(0.25, (0.5, 1.6)).tolist() == [0.5, 0.75, 1.0, 1.25, 1.5]  # stop not on the grid
    assert _saveat_points(1.0, (2.0, 1.0)).tolist() == []  # empty range
    assert _saveat_points([0.0, 5.0], (0.0, 10.0)).tolist() == [0.0, 5.0]
    with pytest.raises(ArgumentError):
        _saveat_points(0.0, (0.0, 1.0))
