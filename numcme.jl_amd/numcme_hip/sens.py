"""Forward sensitivities — mirrors ``src/forwardsensfspmatrix/`` and ``src/forwardsenscme/``.

* ``ForwardSensFspMatrixSparse(model, space)`` (sensfspmatrixsparse.jl:9-95) with
  ``matvec`` / ``matvec_`` (:97-142). Vectors hold ``P+1`` blocks of ``rowcount`` entries:
  ``[p; sinks; S_1; dsinks_1; ...]``.
* ``forwardsens_initial_condition`` / ``ForwardSensFspInitialConditionSparse``
  (forwardsenscmesparse.jl:37-75), ``ForwardSensRStepAdapter``
  (fsspaceadapterssparse.jl:10-14), ``AdaptiveForwardSensFspSparse`` (:16-25) and the
  sensitivity ``solve`` (:99-215) returning ``ForwardSensFspOutputSparse``
  (sensoutputsparse.jl:20-68). The adapt-and-restart loop runs in the library
  (``nc_solve_sens_adaptive``).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from ._lib import ArgumentError
from .fspvector import FspVectorSparse


class ForwardSensFspMatrixSparse(L.Handle):
    """``ForwardSensFspMatrixSparse{RealT}(model, statespace)`` (sensfspmatrixsparse.jl:31-95)."""

    _destroy = "nc_sensmatrix_destroy"

    def __init__(self, model, space, precision="fp64"):
        self.model = model
        h = C.c_void_p()
        prec = L.NC_FP32 if precision == "fp32" else L.NC_FP64
        L.call("nc_sensmatrix_create", space.handle, model.device_model().handle, prec, C.byref(h))
        super().__init__(h.value)
        n, npar, npairs, b = C.c_int64(), C.c_int32(), C.c_int32(), C.c_double()
        L.call("nc_sensmatrix_info", self.handle, C.byref(n), C.byref(npar), C.byref(npairs), C.byref(b))
        self.rowcount, self.parameter_count, self.pair_count = n.value, npar.value, npairs.value
        self.bytes_per_matvec = b.value

    def matvec(self, t, vs):
        """``matvec!`` into a new array (sensfspmatrixsparse.jl:97-142)."""
        v = L.f64(vs)
        if v.size != self.rowcount * (self.parameter_count + 1):
            raise ArgumentError(L.NC_ERR_ARG, "vector length must be (parameter_count+1) * rowcount")
        out = np.empty_like(v)
        L.call("nc_sensmatrix_matvec", self.handle, float(t), L.ptr(v, C.c_double), L.ptr(out, C.c_double))
        return out

    def bench_matvec(self, reps=50, t=0.0):
        ms = C.c_double()
        L.call("nc_sensmatrix_bench_matvec", self.handle, float(t), int(reps), C.byref(ms))
        return ms.value


def matvec_(out, t, SA: ForwardSensFspMatrixSparse, vs):
    out[:] = SA.matvec(t, vs)
    return out


@dataclass
class ForwardSensFspInitialConditionSparse:
    """Initial condition (forwardsenscmesparse.jl:37-53): states, p, S (one vector per parameter)."""

    states: np.ndarray
    p: np.ndarray
    S: list

    def get_states(self):
        return self.states

    def get_probability(self):
        return self.p

    def get_sensitivity(self):
        return self.S


def forwardsens_initial_condition(states, probabilities, sensitivity):
    """``forwardsens_initial_condition`` (forwardsenscmesparse.jl:60-75)."""
    if len(states) == 0:
        raise ArgumentError(L.NC_ERR_ARG, "Empty state list in input.")
    X = np.asarray(states, dtype=np.int64).reshape(len(states), -1)
    return ForwardSensFspInitialConditionSparse(X.copy(), np.asarray(probabilities, dtype=np.float64).copy(),
                                                [np.asarray(s, dtype=np.float64).copy() for s in sensitivity])


@dataclass(frozen=True)
class ForwardSensRStepAdapter:
    """``ForwardSensRStepAdapter`` (fsspaceadapterssparse.jl:10-14)."""

    initial_step_count: int
    max_step_count: int
    dropstates: bool

    def c_struct(self):
        return L.nc_adapter(L.NC_ADAPTER_RSTEP, int(self.initial_step_count), int(self.max_step_count),
                            int(bool(self.dropstates)))


@dataclass
class AdaptiveForwardSensFspSparse:
    """``AdaptiveForwardSensFspSparse`` (forwardsenscmesparse.jl:16-25)."""

    space_adapter: ForwardSensRStepAdapter = None
    ode_method: object = None
    max_adaptations: int = 100000


class ForwardSensFspOutputSliceSparse:
    """``ForwardSensFspOutputSliceSparse`` (sensoutputsparse.jl:33-39)."""

    def __init__(self, t, p, sinks, S, dsinks):
        self.t, self.p, self.sinks, self.S, self.dsinks = t, p, sinks, S, dsinks


class ForwardSensFspOutputSparse:
    """``ForwardSensFspOutputSparse`` (sensoutputsparse.jl:20-26): ``S[i][j]`` is the j-th
    parameter's derivative of ``p[i]``."""

    def __init__(self):
        self.t, self.p, self.sinks, self.S, self.dsinks = [], [], [], [], []
        self.space_ids = []
        self.stats = None

    def __len__(self):
        return len(self.t)

    def __getitem__(self, ind):
        # sensoutputsparse.jl:41-57 (0-based here)
        if isinstance(ind, (list, tuple, np.ndarray)):
            return [self[i] for i in ind]
        if ind >= len(self.t):
            raise ArgumentError(L.NC_ERR_ARG, "Requested index exceeds array limit.")
        return ForwardSensFspOutputSliceSparse(self.t[ind], self.p[ind], self.sinks[ind], self.S[ind], self.dsinks[ind])


def solve_sens(model, initial_condition: ForwardSensFspInitialConditionSparse, tspan,
               alg: AdaptiveForwardSensFspSparse, saveat=(), fsptol=1.0e-6, odeatol=1.0e-10, odertol=1.0e-4,
               verbose=False):
    """Sensitivity ``solve`` (forwardsenscmesparse.jl:99-215)."""
    from .adaptive import _Output, _saveat_points
    from .solve import RK45, KrylovExpv

    npar = model.get_parameter_count()
    if len(initial_condition.S) != npar:
        raise ArgumentError(L.NC_ERR_ARG, "Initial condition does not match CME model. Initial condition must contain "
                                          "`np` sensitivity vectors where `np` is the number of CME model parameters.")
    if not isinstance(alg.space_adapter, ForwardSensRStepAdapter):
        raise ArgumentError(L.NC_ERR_ARG, "space_adapter must be a ForwardSensRStepAdapter")
    o = L.nc_adaptive_opts()
    L.call("nc_adaptive_opts_default", C.byref(o))
    o.adapter = alg.space_adapter.c_struct()
    o.fsptol = float(fsptol)
    o.max_adaptations = int(alg.max_adaptations)
    ode = alg.ode_method
    if ode is None:
        L.call("nc_solve_opts_default", C.byref(o.ode))
        o.ode.method = L.NC_METHOD_AUTO
        o.ode.rtol, o.ode.atol = float(odertol), float(odeatol)
        o.precision = L.NC_FP64
    elif isinstance(ode, (KrylovExpv, RK45)):
        o.ode = ode.opts(odeatol, odertol)
        o.precision = L.NC_FP32 if ode.precision == "fp32" else L.NC_FP64
    else:
        raise ArgumentError(L.NC_ERR_ARG, "ode_method must be KrylovExpv, RK45 or None")
    ns, nr = model.get_species_count(), model.get_reaction_count()
    X = L.i32(np.asarray(initial_condition.states).reshape(-1, ns))
    p0 = L.f64(initial_condition.p)
    n0 = X.shape[0]
    if p0.size != n0 or any(len(s) != n0 for s in initial_condition.S):
        raise ArgumentError(L.NC_ERR_ARG, "initial condition vectors must match the state list")
    S0 = L.f64(np.stack([np.asarray(s, dtype=np.float64) for s in initial_condition.S]))
    ts = _saveat_points(saveat, tspan)
    tsp = L.f64(ts if ts.size else np.zeros(1))
    h = C.c_void_p()
    st = L.nc_solve_stats()
    L.call("nc_solve_sens_adaptive", model.device_model().handle, n0, L.ptr(X, C.c_int32), L.ptr(p0, C.c_double),
           L.ptr(S0, C.c_double), float(tspan[0]), float(tspan[1]), int(ts.size), L.ptr(tsp, C.c_double), C.byref(o),
           C.byref(h), C.byref(st))
    res = _Output(h.value)
    out = ForwardSensFspOutputSparse()
    nsnap, nspace = C.c_int32(), C.c_int32()
    L.call("nc_output_count", res.handle, C.byref(nsnap), C.byref(nspace))
    seen = set()
    for k in range(nsnap.value):
        t, n, sid = C.c_double(), C.c_int64(), C.c_int32()
        L.call("nc_output_info", res.handle, k, C.byref(t), C.byref(n), C.byref(sid))
        nn = n.value
        if sid.value not in seen:
            seen.add(sid.value)
            if verbose and sid.value > 0:
                print(f"At t = {round(t.value, 2)}: update sate space. New size: {nn} states.")
        # every block stays on the device; the snapshots of one space share its state list
        sinks = np.empty(nr)
        h = C.c_void_p()
        L.call("nc_output_get_fspvec", res.handle, k, -1, C.byref(h), L.ptr(sinks, C.c_double))
        p = FspVectorSparse._from_device(h.value)
        Ss, ds = [], []
        for ip in range(npar):
            d = np.empty(nr)
            L.call("nc_output_get_fspvec", res.handle, k, ip, C.byref(h), L.ptr(d, C.c_double))
            Ss.append(FspVectorSparse._from_device(h.value))
            ds.append(d)
        out.t.append(t.value)
        out.space_ids.append(sid.value)
        out.p.append(p)
        out.sinks.append(sinks)
        out.S.append(Ss)
        out.dsinks.append(ds)
    out.stats = {f: getattr(st, f) for f, _ in L.nc_solve_stats._fields_}
    out.stats["adaptations"] = nspace.value - 1
    res.free()
    return out
