"""GPU tests of the Krylov projection sums over 512-row blocks.

Each Arnoldi column of the device expv (the Krylov integrator that replaces the
reference's ODE call around matvec!, fspsolve.jl:10-41 with fspsparsematrix.jl:196-247)
needs V_i . q and V_i . V_j over the rows. The stored generator's mat-vec forms those sums
for its own 512 rows in an epilogue; every other generator (computed modes, more than 8
slots, a mix of computed and loaded columns) runs the mat-vec plainly and a block pass over
the stored q. Both kernels form each block sum from the same values in the same order, so
a solve must give the same snapshots bit for bit whichever ran. Timing mode 2 forces the
plain mat-vec with the block pass on any generator, which gives each generator both forms.
The solves' accuracy against the oracle is tested elsewhere (the C1/C2 parity tests).
"""
import numpy as np
import pytest

import models
import numcme_hip as nh
from numcme_hip.solve import _integrate

pytestmark = pytest.mark.gpu


def _solve(S, states, model, precision, time_kernels, m=30, ts=(0.0, 0.5, 2.0)):
    space = nh.StateSpaceSparse(S, states)
    A = nh.FspMatrixSparse(space, model.propensities, model.parameters, precision=precision, model=model)
    p0 = np.zeros(len(states))
    p0[int(np.flatnonzero((states == 0).all(axis=1))[0])] = 1.0
    u0 = np.concatenate([p0, np.zeros(space.get_sink_count())])
    opts = nh.KrylovExpv(m=m, tol=1e-9, precision=precision, time_kernels=time_kernels).opts(1e-6, 1e-4)
    out = _integrate(A, opts, u0, 0.0, np.array(ts), space)
    return [(p.values, np.asarray(s)) for p, s in zip(out.p, out.sinks)], A


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("order", ["lattice", "shuffled"])
def test_epilogue_equals_block_pass(gpu, precision, order):
    # lattice order: every column computed (DM 1); shuffled: every column loaded (DM 0).
    # 200 x 300 states is not a multiple of 512 rows: the last block is partly past the
    # states, and the 4 sink rows are added by the fold.
    S, props, th = models.toggle()
    states = models.box_states([199, 299])
    if order == "shuffled":
        states = states[np.random.default_rng(3).permutation(len(states))]
    model = nh.CmeModel(S, models.to_product(props), th)
    fused, A = _solve(S, states, model, precision, 0)
    inf = A.info()
    assert inf["rows_per_thread"] == 2 and inf["matrix_free"] == 0
    assert inf["computed_col_slots"] == (4 if order == "lattice" else 0)
    plain, _ = _solve(S, states, model, precision, 2)
    for (pa, sa), (pb, sb) in zip(fused, plain):
        assert np.array_equal(pa, pb) and np.array_equal(sa, sb)
    tot = fused[-1][0].sum() + fused[-1][1].sum()
    assert abs(tot - 1.0) < (1e-6 if precision == "fp32" else 1e-9)


@pytest.mark.parametrize("m", [3, 5, 17, 31])
def test_epilogue_every_basis_count(gpu, m):
    # Krylov dimensions whose columns use 1..m basis vectors, one wave short of a full set of
    # four and beyond (the per-wave vector counts 1..8)
    S, props, th = models.toggle()
    states = models.box_states([63, 63])
    model = nh.CmeModel(S, models.to_product(props), th)
    fused, _ = _solve(S, states, model, "fp32", 0, m=m)
    plain, _ = _solve(S, states, model, "fp32", 2, m=m)
    for (pa, sa), (pb, sb) in zip(fused, plain):
        assert np.array_equal(pa, pb) and np.array_equal(sa, sb)


def test_bench_projection(gpu):
    S, props, th = models.toggle()
    states = models.box_states([127, 127])
    space = nh.StateSpaceSparse(S, states)
    A = nh.FspMatrixSparse(space, models.to_product(props), th, precision="fp32")
    for nv in (1, 4, 5, 30):
        assert A.bench_projection(nv, reps=3) > 0.0
    with pytest.raises(nh.NumCmeHipError):
        A.bench_projection(33, reps=1)
    # the computed-values mode has no projection epilogue
    A.set_matrix_free(1)
    with pytest.raises(nh.NumCmeHipError):
        A.bench_projection(4, reps=1)
    A.set_matrix_free(0)
    assert A.bench_projection(2, reps=1) > 0.0


def test_large_space_grid_stride_pass(gpu):
    # 1500 x 1450 = 2,175,000 states: past the 4,096 blocks of 512 rows, so the Krylov solve
    # takes the grid-stride projection pass; checked against the device DP5(4) solve
    S, props, th = models.toggle()
    states = models.box_states([1499, 1449])
    assert len(states) > 4096 * 512
    model = nh.CmeModel(S, models.to_product(props), th)
    space = nh.StateSpaceSparse(S, states)
    A = nh.FspMatrixSparse(space, model.propensities, model.parameters, precision="fp64", model=model)
    p0 = np.zeros(len(states))
    p0[int(np.flatnonzero((states[:, 0] == 20) & (states[:, 1] == 5))[0])] = 1.0
    u0 = np.concatenate([p0, np.zeros(space.get_sink_count())])
    ts = np.array([0.0, 0.05, 0.1])
    kr = _integrate(A, nh.KrylovExpv(m=30, tol=1e-12).opts(1e-6, 1e-4), u0, 0.0, ts, space)
    rk = _integrate(A, nh.RK45(rtol=1e-10, atol=1e-16).opts(1e-6, 1e-4), u0, 0.0, ts, space)
    for pk, pr, sk, sr in zip(kr.p, rk.p, kr.sinks, rk.sinks):
        assert np.abs(pk.values - pr.values).sum() + np.abs(np.asarray(sk) - np.asarray(sr)).sum() < 1e-8
    assert abs(kr.p[-1].values.sum() + np.sum(kr.sinks[-1]) - 1.0) < 1e-10
