"""Reaction network model — mirrors ``src/cmemodel/`` of NumCME.jl.

* ``StandardTimeInvariantPropensity`` / ``SeparableTimeVaryingPropensity`` /
  ``JointTimeVaryingPropensity`` and ``propensity(...)``: ``propensity.jl:50-153``
* ``CmeModel``: ``cmemodel.jl:49-84``
* ``CmeModelWithSensitivity``: ``cmemodel.jl:104-170``; the propensity gradients and
  the reaction x parameter sparsity pattern are derived in closed form from the traced
  expressions (``numcme_hip.expr.diff``) instead of ForwardDiff/ModelingToolkit.

Indices are 0-based on the Python side: ``x[0]`` is the first species and ``p[0]`` the
first parameter (Julia's ``x[1]``, ``p[1]``).
"""
from __future__ import annotations

import ctypes as C
import inspect

import numpy as np

from . import _lib as L
from . import expr as E
from ._lib import ArgumentError


class Propensity:
    """Base type for all propensities (propensity.jl:8)."""

    def istimevarying(self):
        return False

    def istimeseparable(self):
        return False


class StandardTimeInvariantPropensity(Propensity):
    """``f(x, p)`` (propensity.jl:50-52, call method :145-147)."""

    kind = L.NC_PROP_TIME_INVARIANT

    def __init__(self, f):
        self.f = f

    def __call__(self, x, p=()):
        return self.f(x, p)

    def __repr__(self):
        return f"StandardTimeInvariantPropensity({getattr(self.f, '__name__', self.f)})"


class SeparableTimeVaryingPropensity(Propensity):
    """``tfactor(t, p) * statefactor(x, p)`` (propensity.jl:91-94, :151-153)."""

    kind = L.NC_PROP_SEPARABLE

    def __init__(self, tfactor, statefactor):
        self.tfactor = tfactor
        self.statefactor = statefactor

    def istimevarying(self):
        return True

    def istimeseparable(self):
        return True

    def __call__(self, t, x, p=()):
        return self.tfactor(t, p) * self.statefactor(x, p)


class JointTimeVaryingPropensity(Propensity):
    """``f(t, x, p)`` (propensity.jl:122-124, :148-150)."""

    kind = L.NC_PROP_JOINT

    def __init__(self, f):
        self.f = f

    def istimevarying(self):
        return True

    def __call__(self, t, x, p=()):
        return self.f(t, x, p)


def istimevarying(a: Propensity) -> bool:
    return a.istimevarying()


def istimeseparable(a: Propensity) -> bool:
    return a.istimeseparable()


def propensity(f, tfactor=None):
    """``propensity(f)`` / ``propensity(xfactor, tfactor)`` (propensity.jl:129-142).

    One callable of two arguments ``(x, p)`` gives a time-invariant propensity, one of
    three arguments ``(t, x, p)`` a joint time-varying one; a state factor plus a time
    factor ``tfactor(t, p)`` gives a separable one.
    """
    if tfactor is not None:
        return SeparableTimeVaryingPropensity(tfactor, f)
    try:
        nargs = len(inspect.signature(f).parameters)
    except (TypeError, ValueError):
        nargs = -1
    if nargs == 2:
        return StandardTimeInvariantPropensity(f)
    if nargs == 3:
        return JointTimeVaryingPropensity(f)
    raise ArgumentError(L.NC_ERR_ARG, "The callable passed to `propensity()` must have either two arguments (x,p) "
                                      "or three arguments (t,x,p).")


class _Programs:
    """Traced expressions of a model's propensities and their flattened programs.

    With ``factor_joint`` (the default), a joint law ``f(t, x, p)`` that is a product of a
    time factor and a state factor is lowered as a separable propensity, and one that does
    not read time as a time-invariant one (``expr.factor_time``); ``factored[r]`` records
    the kind reaction ``r`` was lowered from. Other joint laws stay joint.
    """

    def __init__(self, ns, nr, npar, props, factor_joint=True):
        t, x, p = E.symbols(ns, npar)
        self.xexpr, self.texpr, self.kinds, self.factored = [], [], [], []
        for r, a in enumerate(props):
            kind, lowered_from = a.kind if isinstance(a, Propensity) else None, None
            if isinstance(a, StandardTimeInvariantPropensity):
                xe, te = E.trace(a.f, x, p), None
            elif isinstance(a, SeparableTimeVaryingPropensity):
                xe, te = E.trace(a.statefactor, x, p), E.trace(a.tfactor, t, p)
            elif isinstance(a, JointTimeVaryingPropensity):
                xe, te = E.trace(a.f, t, x, p), None
                split = E.factor_time(xe) if factor_joint else None
                if split is not None:
                    lowered_from = L.NC_PROP_JOINT
                    if split[0] == "ti":
                        kind, xe = L.NC_PROP_TIME_INVARIANT, split[1]
                    else:
                        kind, te, xe = L.NC_PROP_SEPARABLE, split[1], split[2]
            else:
                raise ArgumentError(L.NC_ERR_ARG, f"reaction {r}: unsupported propensity type {type(a).__name__}")
            if kind == L.NC_PROP_TIME_INVARIANT and E.uses(xe, "TIME"):
                raise ArgumentError(L.NC_ERR_ARG, f"reaction {r}: time-invariant propensity uses time")
            if te is not None and E.uses(te, "SPECIES"):
                raise ArgumentError(L.NC_ERR_ARG, f"reaction {r}: time factor reads the state")
            self.xexpr.append(xe)
            self.texpr.append(te)
            self.kinds.append(kind)
            self.factored.append(lowered_from)

    @staticmethod
    def flatten(pairs):
        """[(expr_or_None, expr_or_None), ...] -> (offsets int64[2n+1], nc_instr array)."""
        offs, code = [0], []
        for a, b in pairs:
            for e in (a, b):
                if e is not None:
                    code.extend(E.compile_expr(e))
                offs.append(len(code))
        arr = (L.nc_instr * max(len(code), 1))()
        for k, (op, arg, c) in enumerate(code):
            arr[k].op, arr[k].arg, arr[k].c = op, arg, c
        return np.asarray(offs, dtype=np.int64), arr


class CmeModel:
    """Stochastic reaction network model (cmemodel.jl:49-53).

    ``stoich_matrix`` is species x reactions (column r is the net change of reaction r).
    ``factor_joint=False`` keeps every ``JointTimeVaryingPropensity`` joint on the device
    (re-evaluated at every state when ``t`` changes) instead of lowering the factorable
    ones to separable form.
    """

    def __init__(self, stoich_matrix, propensities, parameters=(), factor_joint=True):
        S = np.asarray(stoich_matrix)
        if S.ndim != 2:
            raise ArgumentError(L.NC_ERR_ARG, "stoichiometry matrix must be 2-D (species x reactions)")
        if len(propensities) != S.shape[1]:
            raise ArgumentError(L.NC_ERR_ARG, "need one propensity per reaction (column of the stoichiometry matrix)")
        self.stoich_matrix = S.astype(np.int64)
        self.propensities = list(propensities)
        self.parameters = np.asarray(parameters, dtype=np.float64).ravel()
        self.factor_joint = bool(factor_joint)
        self._progs = None
        self._flat = None
        self._handle = None

    def get_parameters(self):
        return self.parameters

    def get_species_count(self):
        return self.stoich_matrix.shape[0]

    def get_reaction_count(self):
        return self.stoich_matrix.shape[1]

    def get_parameter_count(self):
        return len(self.parameters)

    def get_stoich_matrix(self):
        return self.stoich_matrix

    def get_propensities(self):
        return self.propensities

    def __repr__(self):
        return (f"Stochastic reaction network with {self.get_species_count()} species, "
                f"{self.get_reaction_count()} reactions and {self.get_parameter_count()} parameters.")

    def with_parameters(self, parameters):
        """The same network with other parameter values (a parameter sweep): the traced
        programs do not depend on the values, so they are shared, not traced again."""
        theta = np.asarray(parameters, dtype=np.float64).ravel()
        if len(theta) != self.get_parameter_count():
            raise ArgumentError(L.NC_ERR_ARG, "with_parameters needs as many values as the model has parameters")
        m = CmeModel.__new__(CmeModel)
        m.stoich_matrix, m.propensities, m.factor_joint = self.stoich_matrix, self.propensities, self.factor_joint
        m.parameters = theta
        m._progs = self.programs
        m._flat = self._flatten()
        m._handle = None
        return m

    # ---- device model -----------------------------------------------------------------
    def _flatten(self):
        if getattr(self, "_flat", None) is None:
            pr = self.programs
            self._flat = _Programs.flatten(list(zip(pr.xexpr, pr.texpr)))
        return self._flat

    @property
    def programs(self):
        if self._progs is None:
            self._progs = _Programs(self.get_species_count(), self.get_reaction_count(), self.get_parameter_count(),
                                    self.propensities, self.factor_joint)
        return self._progs

    def device_model(self):
        """The library model handle (created on first use, kept with this object)."""
        if self._handle is None:
            pr = self.programs
            offs, code = self._flatten()
            S = L.i32(self.stoich_matrix.T.copy())  # column-major ns x nr == row-major nr x ns
            theta = L.f64(self.parameters if len(self.parameters) else np.zeros(1))
            kinds = L.i32(pr.kinds)
            h = C.c_void_p()
            L.call("nc_model_create", self.get_species_count(), self.get_reaction_count(), L.ptr(S, C.c_int32),
                   self.get_parameter_count(), L.ptr(theta, C.c_double), L.ptr(kinds, C.c_int32),
                   L.ptr(offs, C.c_int64), code, C.byref(h))
            self._handle = _ModelHandle(h.value)
            self._hold = (offs, code)
        return self._handle

    def eval_propensity(self, r, states, t=0.0, which=0):
        """Evaluate reaction r's x program (which=0) or t program (which=1) on the device."""
        X = L.i32(np.asarray(states).reshape(-1, self.get_species_count()))
        out = np.empty(X.shape[0])
        L.call("nc_model_eval", self.device_model().handle, int(r), int(which), float(t), X.shape[0],
               L.ptr(X, C.c_int32), L.ptr(out, C.c_double))
        return out


class _ModelHandle(L.Handle):
    _destroy = "nc_model_destroy"


class CmeModelWithSensitivity:
    """CME model with propensity derivatives (cmemodel.jl:104-108,157-177)."""

    def __init__(self, model_or_stoich, propensities=None, parameters=None):
        model = model_or_stoich if isinstance(model_or_stoich, CmeModel) else CmeModel(model_or_stoich, propensities,
                                                                                        parameters)
        self.cmemodel = model
        pr = model.programs
        npar = model.get_parameter_count()
        nr = model.get_reaction_count()
        self.gradient_sparsity_patterns = np.zeros((nr, npar), dtype=bool)
        self.dx, self.dt = {}, {}
        for r in range(nr):
            used = E.params_used(pr.xexpr[r])
            if pr.texpr[r] is not None:
                used |= E.params_used(pr.texpr[r])
            for k in sorted(used):
                self.gradient_sparsity_patterns[r, k] = True
                self.dx[(r, k)] = E.diff(pr.xexpr[r], k)
                if pr.texpr[r] is not None:
                    self.dt[(r, k)] = E.diff(pr.texpr[r], k)
        self._grad_set = False

    def get_parameters(self):
        return self.cmemodel.parameters

    def get_species_count(self):
        return self.cmemodel.get_species_count()

    def get_reaction_count(self):
        return self.cmemodel.get_reaction_count()

    def get_parameter_count(self):
        return self.cmemodel.get_parameter_count()

    def get_stoich_matrix(self):
        return self.cmemodel.stoich_matrix

    def get_propensities(self):
        return self.cmemodel.propensities

    def get_gradient_sparsity_patterns(self):
        return self.gradient_sparsity_patterns

    def propensity_gradient(self, r, x, t=0.0):
        """d a_r / d theta at state(s) x — PropensityGradient call semantics (propensitygrad.jl:56-67)."""
        pr = self.cmemodel.programs
        th = self.cmemodel.parameters
        out = []
        for k in range(self.get_parameter_count()):
            if not self.gradient_sparsity_patterns[r, k]:
                out.append(0.0 * np.asarray(x[0], dtype=float))
                continue
            if pr.texpr[r] is None:
                out.append(E.evaluate(self.dx[(r, k)], t, x, th) + 0.0 * np.asarray(x[0], dtype=float))
            else:  # separable: sf * dtf + tf * dsf
                sf = E.evaluate(pr.xexpr[r], t, x, th)
                tf = E.evaluate(pr.texpr[r], t, x, th)
                out.append(sf * E.evaluate(self.dt[(r, k)], t, x, th) + tf * E.evaluate(self.dx[(r, k)], t, x, th))
        return out

    def device_model(self):
        m = self.cmemodel.device_model()
        if not self._grad_set:
            nr, npar = self.get_reaction_count(), self.get_parameter_count()
            pairs = []
            for r in range(nr):
                for k in range(npar):
                    if self.gradient_sparsity_patterns[r, k]:
                        pairs.append((self.dx[(r, k)], self.dt.get((r, k))))
                    else:
                        pairs.append((None, None))
            offs, code = _Programs.flatten(pairs)
            pattern = np.ascontiguousarray(self.gradient_sparsity_patterns.astype(np.uint8))
            L.call("nc_model_set_gradients", m.handle, L.ptr(pattern, C.c_uint8), L.ptr(offs, C.c_int64), code)
            self._grad_hold = (offs, code, pattern)
            self._grad_set = True
        return m
