"""FSP generator — mirrors ``src/fspmatrix/sparse/fspsparsematrix.jl``.

``FspMatrixSparse(space, propensities, parameters=...)`` builds the generator on the GPU
from the device state space (``nc_matrix_create``); ``matvec``/``matvec_``/``matvecadd_``
apply it (``nc_matrix_matvec``). ``precision="fp32"`` stores FP32 off-diagonal values
with an FP64 diagonal that makes every stored column sum to exactly zero.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib as L
from ._lib import ArgumentError
from .model import CmeModel

_PREC = {"fp64": L.NC_FP64, "fp32": L.NC_FP32, np.float64: L.NC_FP64, np.float32: L.NC_FP32}


def _precision(p):
    try:
        return _PREC[p]
    except (KeyError, TypeError):
        raise ArgumentError(L.NC_ERR_ARG, f"unknown precision {p!r}; use 'fp64' or 'fp32'") from None


class FspMatrixSparse(L.Handle):
    """``FspMatrixSparse{RealT}(space, propensities; parameters)`` (fspsparsematrix.jl:47-117)."""

    _destroy = "nc_matrix_destroy"

    def __init__(self, space, propensities, parameters=(), precision="fp64", model=None):
        if model is None:
            model = CmeModel(space.get_stoich_matrix(), propensities, parameters)
        self.model = model  # keeps the device model alive as long as the matrix
        self.space_size = space.get_state_count()
        h = C.c_void_p()
        L.call("nc_matrix_create", space.handle, model.device_model().handle, _precision(precision), C.byref(h))
        super().__init__(h.value)
        self.precision = precision

    def info(self):
        inf = L.nc_matrix_info_t()
        L.call("nc_matrix_info", self.handle, C.byref(inf))
        return {f: getattr(inf, f) for f, _ in L.nc_matrix_info_t._fields_}

    @property
    def rowcount(self):
        return self.info()["n_rows"]

    def size(self, dim=None):
        """``size(A)`` / ``size(A, dim)`` (fspsparsematrix.jl:174-186)."""
        n = self.rowcount
        if dim is None:
            return (n, n)
        if not 1 <= dim <= 2:
            raise ArgumentError(L.NC_ERR_ARG, "Second argument must be either 1 or 2.")
        return n

    def get_parameters(self):
        return self.model.parameters

    def get_propensities(self):
        return self.model.propensities

    def matvec_(self, out, t, v):
        """``matvec!(out, t, A, v)`` (fspsparsematrix.jl:196-217): out = A(t) v."""
        self._mv(out, t, v, 0)
        return out

    def matvecadd_(self, out, t, v):
        """``matvecadd!(out, t, A, v)`` (fspsparsematrix.jl:226-247): out += A(t) v."""
        self._mv(out, t, v, 1)
        return out

    def _mv(self, out, t, v, acc):
        n = self.rowcount
        v = L.f64(v)
        if v.shape != (n,) or not isinstance(out, np.ndarray) or out.shape != (n,) or out.dtype != np.float64:
            raise ArgumentError(L.NC_ERR_ARG, f"matvec needs float64 vectors of length {n}")
        if not out.flags.c_contiguous:
            raise ArgumentError(L.NC_ERR_ARG, "output vector must be contiguous")
        L.call("nc_matrix_matvec", self.handle, float(t), L.ptr(v, C.c_double), L.ptr(out, C.c_double), int(acc))

    def to_scipy(self, t=0.0):
        """The stored generator A(t) as a scipy CSC matrix (duplicates summed)."""
        nnz = C.c_int64()
        L.call("nc_matrix_export", self.handle, float(t), C.byref(nnz), None, None, None)
        rows = np.empty(nnz.value, np.int64)
        cols = np.empty(nnz.value, np.int64)
        vals = np.empty(nnz.value, np.float64)
        L.call("nc_matrix_export", self.handle, float(t), C.byref(nnz), L.ptr(rows, C.c_int64), L.ptr(cols, C.c_int64),
               L.ptr(vals, C.c_double))
        n = self.rowcount
        A = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsc()
        A.sum_duplicates()
        return A

    def set_rows_per_thread(self, rows):
        L.call("nc_matrix_set_rows_per_thread", self.handle, int(rows))

    def set_matrix_free(self, mode):
        """Mat-vec mode: 0 reads the stored generator; 1 takes the off-diagonal values from
        per-reaction tables and the columns from row offsets (box spaces in lattice order,
        time-invariant single-species propensities); 2 computes the diagonal too
        (matrix-free). Results are bit-identical across modes; 0 is the default (the
        fastest on MI355X)."""
        L.call("nc_matrix_set_matrix_free", self.handle, int(mode))

    def autotune(self, reps=50):
        """Pick the fastest mat-vec launch shape for this matrix; returns (rows, ms)."""
        ms = C.c_double()
        L.call("nc_matrix_autotune", self.handle, int(reps), C.byref(ms))
        return self.info()["rows_per_thread"], ms.value

    def bench_matvec(self, reps=100, t=0.0):
        """Mean mat-vec kernel time in ms (HIP events on the matrix stream)."""
        ms = C.c_double()
        L.call("nc_matrix_bench_matvec", self.handle, float(t), int(reps), C.byref(ms))
        return ms.value

    def bench_projection(self, nv, reps=100):
        """Mean time in ms of the Krylov mat-vec with its projection epilogue over nv basis
        vectors (HIP events on the matrix stream)."""
        ms = C.c_double()
        L.call("nc_matrix_bench_projection", self.handle, int(nv), int(reps), C.byref(ms))
        return ms.value


def matvec(t, A: FspMatrixSparse, v):
    """``matvec(t, A, v)`` (fspsparsematrix.jl:254-258)."""
    out = np.zeros(A.rowcount)
    return A.matvec_(out, t, v)


def matvec_(out, t, A, v):
    return A.matvec_(out, t, v)


def matvecadd_(out, t, A, v):
    return A.matvecadd_(out, t, v)
