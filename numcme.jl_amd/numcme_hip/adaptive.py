"""Adaptive FSP solve and R-step space adapters — mirrors ``src/transientcme/sparse/``.

* ``RStepAdapter`` / ``SelectiveRStepAdapter`` (rstepadapters.jl:12-16, 61-65) with
  ``init_`` / ``adapt_`` standing for ``init!`` / ``adapt!`` (:23-52, :74-110). The state
  space is changed in place on the GPU; the probability vector is returned (the reference
  resizes its ``p`` in place, which numpy arrays cannot do).
* ``AdaptiveFspSparse(ode_method, space_adapter)`` (fspsolve.jl:59-62) and the adaptive
  ``solve`` (fspsolve.jl:105-197). The whole adapt-and-restart loop runs in the library
  (``nc_solve_adaptive``): the budget event, the drop selection, ``expand!``, the generator
  rebuild and the restart never leave the device except for a handful of scalars.
"""
from __future__ import annotations

import ctypes as C
import math
from fractions import Fraction
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from ._lib import ArgumentError
from .fspvector import FspVectorSparse


@dataclass(frozen=True)
class RStepAdapter:
    """Reachability adapter (rstepadapters.jl:12-16)."""

    initial_step_count: int
    max_step_count: int
    dropstates: bool

    kind = L.NC_ADAPTER_RSTEP

    def c_struct(self):
        return L.nc_adapter(self.kind, int(self.initial_step_count), int(self.max_step_count), int(bool(self.dropstates)),
                            None, 0.5, 0.0)


@dataclass(frozen=True)
class SelectiveRStepAdapter(RStepAdapter):
    """Reachability adapter over the reactions whose sinks grow (rstepadapters.jl:61-65)."""

    kind = L.NC_ADAPTER_SELECTIVE_RSTEP


@dataclass(frozen=True)
class LearnedAdapter:
    """Learned space adapter (no reference counterpart): ``SelectiveRStepAdapter`` whose
    first expansion level starts only from the boundary states that ``mlp`` (an
    ``AdapterMlp``) keeps, i.e. whose predicted probability is at least ``threshold``.
    Nothing kept, or a sink event sooner than ``fallback_dt`` (0: 1e-3 of the time span)
    after a learned adaptation, makes that adaptation plain ``SelectiveRStepAdapter``. The
    sink bound judges every step exactly as with the plain adapters."""

    mlp: object
    initial_step_count: int
    max_step_count: int
    dropstates: bool = False
    threshold: float = 0.5
    fallback_dt: float = 0.0

    kind = L.NC_ADAPTER_LEARNED

    def c_struct(self):
        return L.nc_adapter(self.kind, int(self.initial_step_count), int(self.max_step_count), int(bool(self.dropstates)),
                            self.mlp.handle, float(self.threshold), float(self.fallback_dt))


def init_(space, adapter: RStepAdapter, p, t=0.0, fsptol=1e-6):
    """``init!`` (rstepadapters.jl:23-28, 74-79); returns p extended with zeros."""
    p = np.asarray(p, dtype=np.float64)
    nold = space.get_state_count()
    ad = adapter.c_struct()
    L.call("nc_adapter_init", space.handle, C.byref(ad))
    return np.concatenate([p, np.zeros(space.get_state_count() - nold)])


def adapt_(space, adapter: RStepAdapter, p, sinks, t, tend, fsptol, dsinks=None):
    """``adapt!`` (rstepadapters.jl:35-52, 86-110).

    ``dsinks`` are the sink entries of ``du = A(t) u`` (what ``get_du!`` gives the
    reference's selective adapter). Returns ``(p_new, dropids)``: the probabilities of the
    adapted space and the sorted 1-based indices that were deleted.
    """
    p = L.f64(p)
    n = p.size
    ad = adapter.c_struct()
    ds = None
    if adapter.kind == L.NC_ADAPTER_SELECTIVE_RSTEP:
        if n == 0:
            raise ArgumentError(L.NC_ERR_ARG, "Empty `p` input in `adapt!`.")
        if dsinks is None:
            raise ArgumentError(L.NC_ERR_ARG, "the selective adapter needs the sink derivatives (dsinks)")
        ds = L.f64(dsinks)
        if ds.size != space.get_sink_count():
            raise ArgumentError(L.NC_ERR_ARG, "dsinks must have one entry per reaction")
    ids = np.empty(max(n, 1), dtype=np.int64)
    m = C.c_int64()
    pp = p if n else np.zeros(1)
    L.call("nc_adapter_adapt", space.handle, C.byref(ad), n, L.ptr(pp, C.c_double),
           None if ds is None else L.ptr(ds, C.c_double), float(t), float(tend), float(fsptol), C.byref(m),
           L.ptr(ids, C.c_int64))
    drop = ids[: m.value].copy()
    keep = np.ones(n, dtype=bool)
    keep[drop] = False
    p_new = np.concatenate([p[keep], np.zeros(space.get_state_count() - int(keep.sum()))])
    return p_new, drop + 1


@dataclass
class AdaptiveFspSparse:
    """``AdaptiveFspSparse`` (fspsolve.jl:59-62). ``ode_method`` is ``KrylovExpv``,
    ``RK45`` or None (Krylov for time-invariant models, RK45 otherwise)."""

    ode_method: object = None
    space_adapter: RStepAdapter = None
    max_adaptations: int = 100000


def _saveat_points(saveat, tspan):
    if isinstance(saveat, (int, float)) and not isinstance(saveat, bool):
        # Julia's tspan[1]:saveat:tspan[2]. Julia evaluates a float range through exact
        # rationals when start, step and stop have short ones (0:0.1:1 holds 0.3, not
        # 3 * 0.1): here the shortest decimal of each float, used when it converts back
        # exactly, gives elements start + i*step correctly rounded from the rationals.
        a, b, h = float(tspan[0]), float(tspan[1]), float(saveat)
        if h <= 0:
            raise ArgumentError(L.NC_ERR_ARG, "saveat step must be positive")
        fa, fb, fh = (Fraction(repr(x)) for x in (a, b, h))
        if float(fa) == a and float(fb) == b and float(fh) == h and max(fa.denominator, fh.denominator) <= 10**9:
            k = math.floor((fb - fa) / fh)
            return np.array([float(fa + i * fh) for i in range(k + 1)], dtype=np.float64)
        k = math.floor((b - a) / h * (1 + 1e-12) + 1e-12)
        return np.array([a + i * h for i in range(k + 1)], dtype=np.float64)
    return np.asarray(list(saveat), dtype=np.float64)


def _adaptive_opts(alg, fsptol, odeatol, odertol, verbose, records):
    from .solve import RK45, KrylovExpv

    if not isinstance(alg.space_adapter, (RStepAdapter, LearnedAdapter)):
        raise ArgumentError(L.NC_ERR_ARG,
                            "space_adapter must be RStepAdapter, SelectiveRStepAdapter or LearnedAdapter")
    o = L.nc_adaptive_opts()
    L.call("nc_adaptive_opts_default", C.byref(o))
    o.adapter = alg.space_adapter.c_struct()
    o.verbose = 1 if verbose else 0
    o.records = None if records is None else records.handle
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
    return o


def _initial(model, initial_distribution):
    X = L.i32(np.asarray(initial_distribution.states).reshape(-1, model.get_species_count()))
    p0 = L.f64(initial_distribution.values)
    if X.shape[0] != p0.size:
        raise ArgumentError(L.NC_ERR_ARG, "initial distribution states and values differ in length")
    return X, p0


def _output(res, model, st):
    """FspOutputSparse of an nc_output (an ``_Output``, freed before returning); the
    snapshots stay on the device, and those of one space share its state list."""
    from .solve import FspOutputSparse

    out = FspOutputSparse()
    out.space_ids = []  # index of the state space each snapshot lives on
    nsnap, nspace = C.c_int32(), C.c_int32()
    L.call("nc_output_count", res.handle, C.byref(nsnap), C.byref(nspace))
    nr = model.get_reaction_count()
    for k in range(nsnap.value):
        t, n, sid = C.c_double(), C.c_int64(), C.c_int32()
        L.call("nc_output_info", res.handle, k, C.byref(t), C.byref(n), C.byref(sid))
        sinks = np.empty(nr)
        h = C.c_void_p()
        L.call("nc_output_get_fspvec", res.handle, k, -1, C.byref(h), L.ptr(sinks, C.c_double))
        out.t.append(t.value)
        out.space_ids.append(sid.value)
        out.p.append(FspVectorSparse._from_device(h.value))
        out.sinks.append(sinks)
    out.stats = {f: getattr(st, f) for f, _ in L.nc_solve_stats._fields_}
    res.free()
    return out


def solve_adaptive(model, initial_distribution: FspVectorSparse, tspan, alg: AdaptiveFspSparse, saveat=(),
                   fsptol=1.0e-6, odeatol=1.0e-6, odertol=1.0e-4, verbose=False, records=None):
    """The adaptive solve. ``verbose`` prints every space update as the reference does;
    ``records`` (a ``TrainingRecords``) collects learned-adapter training records."""
    o = _adaptive_opts(alg, fsptol, odeatol, odertol, verbose, records)
    ts = _saveat_points(saveat, tspan)
    X, p0 = _initial(model, initial_distribution)
    h = C.c_void_p()
    st = L.nc_solve_stats()
    tsp = ts if ts.size else np.zeros(1)
    L.call("nc_solve_adaptive", model.device_model().handle, X.shape[0], L.ptr(X, C.c_int32), L.ptr(p0, C.c_double),
           float(tspan[0]), float(tspan[1]), int(ts.size), L.ptr(L.f64(tsp), C.c_double), C.byref(o), C.byref(h),
           C.byref(st))
    return _output(_Output(h.value), model, st)


def solve_adaptive_batch(models, initial_distributions, tspan, alg: AdaptiveFspSparse, saveat=(), fsptol=1.0e-6,
                         odeatol=1.0e-6, odertol=1.0e-4, workers=0):
    """Adaptive solves of many instances at once (a parameter sweep; no reference
    counterpart, where a sweep is a loop of ``solve`` calls): ``models[i]`` from
    ``initial_distributions[i]`` (one distribution for all, if a single FspVectorSparse is
    given), all over the same ``tspan``, ``saveat`` and ``alg``. ``workers`` host threads of
    the library (0: 4) each run whole solves on streams of their own. Returns one
    ``FspOutputSparse`` per instance, each what ``solve`` gives for it."""
    models = list(models)
    n = len(models)
    if isinstance(initial_distributions, FspVectorSparse):
        initial_distributions = [initial_distributions] * n
    initial_distributions = list(initial_distributions)
    if len(initial_distributions) != n:
        raise ArgumentError(L.NC_ERR_ARG, "one initial distribution per model")
    o = _adaptive_opts(alg, fsptol, odeatol, odertol, False, None)
    ts = _saveat_points(saveat, tspan)
    if n == 0:
        return []
    parts = [_initial(m, d) for m, d in zip(models, initial_distributions)]
    X = np.ascontiguousarray(np.concatenate([x.ravel() for x, _ in parts]), dtype=np.int32)
    p0 = np.ascontiguousarray(np.concatenate([p for _, p in parts]), dtype=np.float64)
    n0 = np.array([p.size for _, p in parts], dtype=np.int64)
    handles = (C.c_void_p * n)(*[m.device_model().handle for m in models])
    outs = (C.c_void_p * n)()
    stats = (L.nc_solve_stats * n)()
    tsp = ts if ts.size else np.zeros(1)
    L.call("nc_solve_adaptive_batch", n, handles, L.ptr(n0, C.c_int64), L.ptr(X, C.c_int32), L.ptr(p0, C.c_double),
           float(tspan[0]), float(tspan[1]), int(ts.size), L.ptr(L.f64(tsp), C.c_double), C.byref(o), int(workers),
           outs, stats)
    owned = [_Output(outs[i]) for i in range(n)]  # freed even if a conversion fails
    return [_output(res, m, stats[i]) for i, (res, m) in enumerate(zip(owned, models))]


class _Output(L.Handle):
    _destroy = "nc_output_destroy"
