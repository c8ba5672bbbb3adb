"""The allocation-free mat-vec (nc_matrix_matvec_noalloc) against nc_matrix_matvec, through
the C ABI: two calls in a row, the second on a smaller generator, in FP64 and FP32, for
matvec and matvecadd (time-invariant and time-varying generators), and the size limit of
its fixed scratch array. Both entry points run the same kernels, so their results must be
identical bit for bit."""
import ctypes as C

import numpy as np
import pytest

import models
import numcme_hip as nh
from numcme_hip import _lib as L

pytestmark = pytest.mark.gpu


def _both(A, t, v, base=None):
    n = A.rowcount
    v = L.f64(v)
    outs = []
    for name in ("nc_matrix_matvec_noalloc", "nc_matrix_matvec"):
        out = np.zeros(n) if base is None else base.copy()
        L.call(name, A.handle, float(t), L.ptr(v, C.c_double), L.ptr(out, C.c_double), 0 if base is None else 1)
        outs.append(out)
    return outs


def _toggle(precision, side):
    S, props, th = models.toggle()
    sp = nh.StateSpaceSparse(S, models.box_states([side - 1, side - 1]))
    return nh.FspMatrixSparse(sp, models.to_product(props), th, precision=precision)


def _telegraph_tv(precision):
    S, props, th = models.telegraph_tv("joint")
    sp = nh.StateSpaceSparse(S, models.telegraph_states(799))
    model = nh.CmeModel(S, models.to_product(props), th, factor_joint=False)
    return nh.FspMatrixSparse(sp, None, precision=precision, model=model)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_matches_matvec_then_smaller(gpu, precision):
    rng = np.random.default_rng(12)
    big = _toggle(precision, 700)       # 490,000 states
    small = _telegraph_tv(precision)    # 1,600 states, joint time-varying slot
    for A, t in ((big, 0.0), (small, 0.7), (big, 0.0), (small, 2.3)):
        v = rng.random(A.rowcount)
        a, b = _both(A, t, v)
        assert np.array_equal(a, b), (A.rowcount, t)
        assert np.abs(a).max() > 0
        base = rng.random(A.rowcount)
        a, b = _both(A, t, v, base)
        assert np.array_equal(a, b), (A.rowcount, t, "accumulate")


def test_noalloc_rejects_generators_beyond_its_scratch(gpu):
    A = _toggle("fp64", 1800)  # 3,240,004 rows need more than the 48 MiB scratch
    v = np.ones(A.rowcount)
    out = np.zeros(A.rowcount)
    with pytest.raises(nh.ArgumentError, match="scratch"):
        L.call("nc_matrix_matvec_noalloc", A.handle, 0.0, L.ptr(v, C.c_double), L.ptr(out, C.c_double), 0)
    # the library stays usable afterwards
    small = _toggle("fp64", 50)
    a, b = _both(small, 0.0, np.random.default_rng(3).random(small.rowcount))
    assert np.array_equal(a, b)
