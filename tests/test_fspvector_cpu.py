"""FspVectorSparse host logic: the cases of test/test_fspvec.jl, and the marginal sums and
dense conversion of src/fspvector/fspvector.jl:57-129 checked by hand-computed values.
No device call is made."""
import numpy as np
import pytest

import numcme_hip as nh
from numcme_hip._lib import ArgumentError


def _reference_vector():
    # test_fspvec.jl:5-7: states (0, 0..10) with value 0 and (1, 0..10) with value 1
    states = [[0, i] for i in range(11)] + [[1, i] for i in range(11)]
    vals = [0.0] * 11 + [1.0] * 11
    return nh.FspVectorSparse(states, vals)


def test_reference_fspvec_cases():
    v = _reference_vector()
    v1, v2 = v.sum([1]), v.sum([2])
    assert v.sum() == 11.0  # test_fspvec.jl:12
    assert nh.nnz(v1) == 11 and nh.nnz(v2) == 2  # :13-14
    with pytest.raises(ArgumentError):  # :16
        nh.FspVectorSparse([[0, 1]], [1.0, 2.0])
    with pytest.raises(ArgumentError):  # :17
        nh.FspVectorSparse([[0, 1], [1, 2]], [1.0])


def test_marginals_by_hand():
    v = _reference_vector()
    # summing out species 1 leaves species 2's counts 0..10, each with 0 + 1
    v1 = v.sum([1])
    assert v1.states[:, 0].tolist() == list(range(11))
    assert v1.values.tolist() == [1.0] * 11
    # summing out species 2 leaves species 1 in order of first appearance: 0, then 1
    v2 = v.sum([2])
    assert v2.states[:, 0].tolist() == [0, 1] and v2.values.tolist() == [0.0, 11.0]
    # summing out both leaves one empty-coordinate state with the total
    v12 = v.sum([1, 2])
    assert v12.states.shape == (1, 0) and v12.values.tolist() == [11.0]
    # a repeated dimension counts once
    assert v.sum([2, 2]).values.tolist() == [0.0, 11.0]


def test_marginal_first_appearance_order():
    p = nh.FspVectorSparse([[3, 0], [1, 5], [3, 2], [0, 1], [1, 1]], [0.1, 0.2, 0.3, 0.15, 0.25])
    m = p.sum([2])
    assert m.states[:, 0].tolist() == [3, 1, 0]
    assert np.allclose(m.values, [0.4, 0.45, 0.15], rtol=0, atol=1e-15)


def test_dimension_and_empty_errors():
    v = _reference_vector()
    with pytest.raises(ArgumentError):
        v.sum([0])
    with pytest.raises(ArgumentError):
        v.sum([3])
    with pytest.raises(ArgumentError):
        nh.FspVectorSparse(np.zeros((0, 2), dtype=np.int64), []).to_array()


def test_dense_array():
    p = nh.FspVectorSparse([[2, 0], [0, 1], [1, 1]], [0.5, 0.25, 0.25])
    a = p.to_array()
    assert a.shape == (3, 2)
    expect = np.zeros((3, 2))
    expect[2, 0], expect[0, 1], expect[1, 1] = 0.5, 0.25, 0.25
    assert np.array_equal(a, expect)
    assert p.state2idx[(0, 1)] == 2  # 1-based, as the reference's state dictionary


def test_output_indexing():
    """FspOutputSparse indexing (fspoutput.jl:195-209, 0-based here): slices, index lists,
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
