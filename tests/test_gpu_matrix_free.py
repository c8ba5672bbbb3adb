"""GPU tests of the computed mat-vecs (``nc_matrix_set_matrix_free``).

On a box state space in lattice order whose propensities are time-invariant and each read
one species, the mat-vec can take its off-diagonal values from per-reaction tables and its
columns from row offsets (mode 1, the stored FP64 diagonal still read) or compute the
diagonal too (mode 2, matrix-free), instead of loading the stored generator
(fspsparsematrix.jl:47-108 builds and stores it; matvec!/matvecadd!, :196-247, read it).
The tables hold the build's own propensity values and the kernels follow the build's
arithmetic, so every mode must agree with the stored one bit for bit; the stored mode is
itself checked against the oracle elsewhere, and here once more directly.
"""
import numpy as np
import pytest

import models
import numcme_hip as nh
from oracle.fspmatrix import FspMatrix
from oracle.statespace import StateSpace

pytestmark = pytest.mark.gpu


def _modes(A, t, v, modes, accumulate=False, out0=None):
    """out of every mode, the stored mode (0) last; the matrix's default mode is restored"""
    default = A.info()["matrix_free"]
    res = []
    for mode in list(modes) + [0]:
        A.set_matrix_free(mode)
        assert A.info()["matrix_free"] == mode
        out = np.array(out0, dtype=np.float64) if out0 is not None else np.zeros(A.rowcount)
        (A.matvecadd_ if accumulate else A.matvec_)(out, t, v)
        res.append(out)
    A.set_matrix_free(default)
    return res


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("bounds", [(299, 199), (1023, 1023)])
def test_toggle_box_bit_identical_and_vs_oracle(gpu, precision, bounds):
    S, props, th = models.toggle()
    states = models.box_states(list(bounds))
    space = nh.StateSpaceSparse(S, states)
    A = nh.FspMatrixSparse(space, models.to_product(props), th, precision=precision)
    inf = A.info()
    assert inf["matrix_free_available"] == 3 and inf["matrix_free"] == 0
    A.set_matrix_free(1)
    assert A.info()["bytes_per_matvec"] < 0.5 * inf["csr_bytes_per_matvec"]
    A.set_matrix_free(2)
    # what is left is reading v and writing out (16 of the CSR form's ~80 bytes a row in FP64)
    assert A.info()["bytes_per_matvec"] < 0.25 * inf["csr_bytes_per_matvec"]
    A.set_matrix_free(0)
    rng = np.random.default_rng(len(states))
    v = rng.standard_normal(A.rowcount)
    m1, m2, stored = _modes(A, 0.0, v, (1, 2))
    assert np.array_equal(m1, stored) and np.array_equal(m2, stored)
    mf = m1
    out0 = rng.standard_normal(A.rowcount)
    a1, a2, storeda = _modes(A, 0.0, v, (1, 2), accumulate=True, out0=out0)
    assert np.array_equal(a1, storeda) and np.array_equal(a2, storeda)
    if len(states) <= 100_000:
        ref = FspMatrix(StateSpace(S, states), models.to_oracle(props), th).to_csc(0.0) @ v
        tol = 1e-12 if precision == "fp64" else 2e-6
        assert np.max(np.abs(mf - ref)) <= tol * np.max(np.abs(ref))


def _mixed_model():
    # three species; reactions 0 and 1 share a stoichiometry (one slot, two members);
    # reaction 2 has a constant propensity and removes species 2 (a leak at its lower
    # bound); Hill and mass-action factors; the box does not start at zero
    S = np.array([[1, 1, -1, 0, 0],
                  [0, 0, 0, 1, -1],
                  [0, 0, -1, 0, 1]])
    props = [
        ("ti", lambda x, p: p[0] / (1.0 + (x[1] / p[1]) ** 2)),
        ("ti", lambda x, p: p[2] * x[1]),
        ("ti", lambda x, p: p[3] + 0.0 * p[0]),
        ("ti", lambda x, p: p[4] * nh.nmath.sqrt(x[2] + 1.0)),
        ("ti", lambda x, p: p[5] * x[1] * (x[1] - 1.0)),
    ]
    return S, props, np.array([4.0, 3.0, 0.7, 0.9, 1.3, 0.2])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_mixed_model_offset_box(gpu, precision):
    S, props, th = _mixed_model()
    lo, hi = np.array([2, 1, 0]), np.array([13, 9, 6])
    states = models.box_states(list(hi - lo)) + lo
    space = nh.StateSpaceSparse(S, states)
    A = nh.FspMatrixSparse(space, models.to_product(props), th, precision=precision)
    # two reactions share a stoichiometry: only the matrix-free form (mode 2) applies
    inf = A.info()
    assert inf["matrix_free_available"] == 2 and inf["matrix_free"] == 0
    with pytest.raises(nh.NumCmeHipError):
        A.set_matrix_free(1)
    rng = np.random.default_rng(5)
    for t in (0.0, 2.5):
        v = rng.random(A.rowcount)
        mf, stored = _modes(A, t, v, (2,))
        assert np.array_equal(mf, stored)
    ref = FspMatrix(StateSpace(S, states), models.to_oracle(props), th).to_csc(0.0) @ v
    tol = 1e-12 if precision == "fp64" else 2e-6
    assert np.max(np.abs(mf - ref)) <= tol * np.max(np.abs(ref))


def test_ineligible_generators_keep_the_stored_arrays(gpu):
    S, props, th = models.toggle()
    states = models.box_states([40, 30])
    # a propensity that reads two species
    two = list(props)
    two[1] = ("ti", lambda x, p: p[4] * x[0] * (1.0 + 0.01 * x[1]))
    cases = [
        (states, two),
        (states[np.random.default_rng(0).permutation(len(states))], props),  # not in lattice order
        (states[:-1], props),  # not a full box
    ]
    for st, pr in cases:
        A = nh.FspMatrixSparse(nh.StateSpaceSparse(S, st), models.to_product(pr), th)
        inf = A.info()
        assert inf["matrix_free_available"] == 0 and inf["matrix_free"] == 0
        for mode in (1, 2):
            with pytest.raises(nh.NumCmeHipError):
                A.set_matrix_free(mode)
        A.set_matrix_free(0)  # always allowed
    # time-varying propensities
    Stv, props_tv, th_tv = models.telegraph_tv("sep")
    st = models.box_states([1, 1, 20])
    A = nh.FspMatrixSparse(nh.StateSpaceSparse(Stv, st), models.to_product(props_tv), th_tv)
    assert A.info()["matrix_free_available"] == 0


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_krylov_solve_bit_identical(gpu, precision):
    # the toggle switch on 512^2 states: the whole Krylov solve (graph-replayed Arnoldi)
    # gives the same snapshots in every mode
    S, props, th = models.toggle()
    states = models.box_states([511, 511])
    model = nh.CmeModel(S, models.to_product(props), th)
    p0 = np.zeros(len(states))
    p0[0] = 1.0
    from numcme_hip.solve import _integrate

    outs = []
    for mode in (1, 2, 0):
        space = nh.StateSpaceSparse(S, states)
        A = nh.FspMatrixSparse(space, model.propensities, model.parameters, precision=precision, model=model)
        A.set_matrix_free(mode)
        u0 = np.concatenate([p0, np.zeros(space.get_sink_count())])
        out = _integrate(A, nh.KrylovExpv(m=30, tol=1e-9, precision=precision).opts(1e-6, 1e-4), u0, 0.0,
                         np.array([0.0, 1.0, 3.0]), space)
        outs.append([(p.values, s) for p, s in zip(out.p, out.sinks)])
    for other in outs[:2]:
        for (pa, sa), (pb, sb) in zip(other, outs[2]):
            assert np.array_equal(pa, pb) and np.array_equal(sa, sb)
    assert abs(outs[0][-1][0].sum() + outs[0][-1][1].sum() - 1.0) < 1e-6
