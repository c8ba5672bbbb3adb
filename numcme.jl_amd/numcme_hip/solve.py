"""Transient CME solves — mirrors ``src/transientcme/`` of NumCME.jl.

* ``solve(model, p0, tspan, method; saveat, ...)`` with an integrator object
  (``KrylovExpv`` or ``RK45``) is the fixed-space solve of ``fspsolve.jl:10-41``: the
  state space is ``p0``'s state list in order, the ODE state is ``[p0; zeros(R)]`` and
  one snapshot is returned per saved time. The integration runs on the GPU
  (``nc_solver_*``).
* ``FspOutputSparse`` / ``FspOutputSliceSparse``: ``fspoutput.jl:19-74``.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from ._lib import ArgumentError
from .fspmatrix import FspMatrixSparse
from .fspvector import FspVectorSparse
from .model import CmeModel
from .statespace import StateSpaceSparse


@dataclass
class KrylovExpv:
    """Device Krylov ``expv`` integrator for time-invariant generators."""

    m: int = 30
    tol: float = 1e-12
    precision: str = "fp64"
    # HIP events around every mat-vec launch: 1 (or True) as the solve runs them (the
    # mat-vec with its projection epilogue where the generator has one), 2 plain mat-vecs
    # (the projection sums then run as a separate pass; same results)
    time_kernels: int = 0
    graphs: bool = True  # replay each Arnoldi process as a captured HIP graph

    def opts(self, odeatol, odertol):
        o = L.nc_solve_opts()
        L.call("nc_solve_opts_default", C.byref(o))
        o.method = L.NC_METHOD_KRYLOV
        o.krylov_m = self.m
        o.krylov_tol = self.tol
        o.time_kernels = int(self.time_kernels)
        o.no_graphs = 0 if self.graphs else 1
        return o


@dataclass
class RK45:
    """Device Dormand-Prince 5(4); tolerances default to the ``solve`` keywords."""

    rtol: float | None = None
    atol: float | None = None
    precision: str = "fp64"
    time_kernels: bool = False

    def opts(self, odeatol, odertol):
        o = L.nc_solve_opts()
        L.call("nc_solve_opts_default", C.byref(o))
        o.method = L.NC_METHOD_RK45
        o.rtol = self.rtol if self.rtol is not None else odertol
        o.atol = self.atol if self.atol is not None else odeatol
        o.time_kernels = int(self.time_kernels)
        return o


class FspOutputSliceSparse:
    """``FspOutputSliceSparse`` (fspoutput.jl:30-34): fields ``t``, ``p``, ``sinks``."""

    def __init__(self, t, p, sinks):
        self.t, self.p, self.sinks = t, p, sinks


class FspOutputSparse:
    """``FspOutputSparse`` (fspoutput.jl:19-23): lists ``t``, ``p``, ``sinks``."""

    def __init__(self):
        self.t, self.p, self.sinks = [], [], []
        self.stats = None

    def __len__(self):
        return len(self.t)

    def __getitem__(self, ind):
        # fspoutput.jl:44-58 (0-based here)
        if isinstance(ind, (list, tuple, np.ndarray)):
            return [self[i] for i in ind]
        if ind >= len(self.t):
            raise ArgumentError(L.NC_ERR_ARG, "Requested index exceeds array limit.")
        return FspOutputSliceSparse(self.t[ind], self.p[ind], self.sinks[ind])


def solve(model: CmeModel, initial_distribution: FspVectorSparse, tspan, method, saveat=(), fsptol=1.0e-6,
          odeatol=1.0e-6, odertol=1.0e-4, verbose=False, records=None):
    """FSP solve. ``method`` is an integrator (fixed space, fspsolve.jl:10-41) or an
    ``AdaptiveFspSparse`` (adaptive space, fspsolve.jl:105-197). ``records`` (adaptive
    solves only) collects learned-adapter training records."""
    from .adaptive import AdaptiveFspSparse, solve_adaptive
    from .sens import AdaptiveForwardSensFspSparse, solve_sens

    if isinstance(method, AdaptiveForwardSensFspSparse):
        # solve(model::CmeModelWithSensitivity, initial_condition, tspan, alg) (forwardsenscmesparse.jl:99)
        return solve_sens(model, initial_distribution, tspan, method, saveat=saveat, fsptol=fsptol,
                          odeatol=odeatol, odertol=odertol, verbose=verbose)
    if isinstance(method, AdaptiveFspSparse):
        return solve_adaptive(model, initial_distribution, tspan, method, saveat=saveat, fsptol=fsptol,
                              odeatol=odeatol, odertol=odertol, verbose=verbose, records=records)
    if records is not None:
        raise ArgumentError(L.NC_ERR_ARG, "training records come from adaptive solves")
    return solve_fixed(model, initial_distribution, tspan, method, saveat=saveat, odeatol=odeatol, odertol=odertol)


def solve_fixed(model, p0, tspan, method, saveat=(), odeatol=1.0e-6, odertol=1.0e-4):
    if not isinstance(method, (KrylovExpv, RK45)):
        raise ArgumentError(L.NC_ERR_ARG, "method must be KrylovExpv, RK45 or AdaptiveFspSparse")
    space = StateSpaceSparse(model.stoich_matrix, p0.states)  # fspsolve.jl:16
    S = space.get_state_count()
    R = space.get_sink_count()
    if S != len(p0.values):
        raise ArgumentError(L.NC_ERR_ARG, "initial distribution has duplicate or negative states")
    A = FspMatrixSparse(space, model.propensities, model.parameters, precision=method.precision, model=model)
    t0, t1 = float(tspan[0]), float(tspan[1])
    ts = np.asarray(list(saveat), dtype=np.float64) if len(saveat) else np.asarray([t0, t1])
    if np.any(ts < t0) or np.any(ts > t1) or np.any(np.diff(ts) < 0):
        raise ArgumentError(L.NC_ERR_ARG, "saveat must be ascending and inside tspan")
    u0 = np.concatenate([p0.values, np.zeros(R)])  # fspsolve.jl:21
    return _integrate(A, method.opts(odeatol, odertol), u0, t0, ts, space)


def _integrate(A, opts, u0, t0, ts, space):
    """Run the solver and hand out every snapshot as a device-resident vector."""
    h = C.c_void_p()
    L.call("nc_solver_create", A.handle, C.byref(opts), C.byref(h))
    solver = _Solver(h.value)
    u0 = L.f64(u0)
    L.call("nc_solver_set_state", solver.handle, t0, L.ptr(u0, C.c_double))
    tsa = L.f64(ts)
    L.call("nc_solver_run", solver.handle, len(tsa), L.ptr(tsa, C.c_double))
    out = FspOutputSparse()
    R = space.get_sink_count()
    for k, t in enumerate(tsa):
        sinks = np.empty(R)
        h = C.c_void_p()
        L.call("nc_solver_get_fspvec", solver.handle, k, 0, space.handle, C.byref(h), L.ptr(sinks, C.c_double))
        out.t.append(float(t))
        out.p.append(FspVectorSparse._from_device(h.value))
        out.sinks.append(sinks)
    st = L.nc_solve_stats()
    L.call("nc_solver_stats", solver.handle, C.byref(st))
    out.stats = {f: getattr(st, f) for f, _ in L.nc_solve_stats._fields_}
    solver.free()
    return out


class _Solver(L.Handle):
    _destroy = "nc_solver_destroy"
