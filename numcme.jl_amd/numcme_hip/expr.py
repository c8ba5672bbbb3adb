"""Lowering of propensity callables to device programs.

NumCME.jl propensities are Julia closures called as ``f(x, p)``, ``f(t, x, p)`` or
``tfactor(t, p)`` (``src/cmemodel/propensity.jl:145-153``). Closures cannot run on the
GPU, so the host traces them: the callable is called once with symbolic arguments —
``x[s]`` is species ``s``, ``p[k]`` parameter ``k`` (both 0-based), ``t`` time — and the
resulting expression tree is compiled to the postfix programs of the C ABI
(``nc_instr``). Arithmetic operators work directly; other functions come from this
module (``exp``, ``log``, ``sin``, ``cos``, ``sqrt``, ``abs``, ``max``, ``min``,
``ifelse``), which fall back to numpy for numeric arguments, so one callable serves the
device and a numpy check alike.

The same trees give closed-form parameter derivatives (``diff``) and the
reaction x parameter dependency pattern that NumCME.jl obtains from ForwardDiff and
ModelingToolkit (``src/cmemodel/senstools/forwarddiffpropensity.jl:3-37``,
``sparsity_pattern.jl:3-31``). ``max``/``min`` follow Julia's ``Base.max``:
``max(a, b) = b < a ? a : b``, and their derivative follows the selected branch, as
ForwardDiff's dual numbers do.
"""
from __future__ import annotations

import math
import numbers

import numpy as np

# opcodes of include/numcme_hip.h
OP = dict(CONST=0, PARAM=1, SPECIES=2, TIME=3, ADD=4, SUB=5, MUL=6, DIV=7, POW=8, NEG=9, EXP=10, LOG=11,
          SIN=12, COS=13, SQRT=14, ABS=15, MAX=16, MIN=17, LT=18, LE=19, GT=20, GE=21, SELECT=22)
_UNARY = {"NEG", "EXP", "LOG", "SIN", "COS", "SQRT", "ABS"}
_BINARY = {"ADD", "SUB", "MUL", "DIV", "POW", "MAX", "MIN", "LT", "LE", "GT", "GE"}
STACK_LIMIT = 16

pi = math.pi


class Expr:
    """Node of a traced propensity expression."""

    __slots__ = ("op", "args", "value")
    __array_priority__ = 1000  # numpy scalars defer to Expr operators

    def __init__(self, op, args=(), value=None):
        self.op = op
        self.args = tuple(args)
        self.value = value

    # -- construction helpers -------------------------------------------------------
    @staticmethod
    def wrap(v):
        if isinstance(v, Expr):
            return v
        if isinstance(v, (numbers.Real, np.number, bool, np.bool_)):
            return Expr("CONST", value=float(v))
        raise TypeError(f"cannot use {type(v).__name__} in a propensity expression")

    def __add__(self, o): return _bin("ADD", self, o)
    def __radd__(self, o): return _bin("ADD", o, self)
    def __sub__(self, o): return _bin("SUB", self, o)
    def __rsub__(self, o): return _bin("SUB", o, self)
    def __mul__(self, o): return _bin("MUL", self, o)
    def __rmul__(self, o): return _bin("MUL", o, self)
    def __truediv__(self, o): return _bin("DIV", self, o)
    def __rtruediv__(self, o): return _bin("DIV", o, self)
    def __pow__(self, o): return _bin("POW", self, o)
    def __rpow__(self, o): return _bin("POW", o, self)
    def __neg__(self): return _un("NEG", self)
    def __pos__(self): return self
    def __abs__(self): return _un("ABS", self)
    def __lt__(self, o): return _bin("LT", self, o)
    def __le__(self, o): return _bin("LE", self, o)
    def __gt__(self, o): return _bin("GT", self, o)
    def __ge__(self, o): return _bin("GE", self, o)

    def __bool__(self):
        raise TypeError("propensity expressions cannot drive Python control flow; use numcme_hip.expr.ifelse")

    def __repr__(self):
        if self.op == "CONST":
            return repr(self.value)
        if self.op in ("PARAM", "SPECIES"):
            return f"{'p' if self.op == 'PARAM' else 'x'}[{self.value}]"
        if self.op == "TIME":
            return "t"
        return f"{self.op.lower()}({', '.join(map(repr, self.args))})"


def _is_const(e, v=None):
    return e.op == "CONST" and (v is None or e.value == v)


def _bin(op, a, b):
    a, b = Expr.wrap(a), Expr.wrap(b)
    if _is_const(a) and _is_const(b):
        return Expr("CONST", value=float(_eval_node(op, [a.value, b.value])))
    # algebraic identities that are exact in IEEE arithmetic for finite operands
    if op == "ADD":
        if _is_const(a, 0.0):
            return b
        if _is_const(b, 0.0):
            return a
    if op == "SUB" and _is_const(b, 0.0):
        return a
    if op == "MUL":
        if _is_const(a, 1.0):
            return b
        if _is_const(b, 1.0):
            return a
    if op == "DIV" and _is_const(b, 1.0):
        return a
    return Expr(op, (a, b))


def _un(op, a):
    a = Expr.wrap(a)
    if _is_const(a):
        return Expr("CONST", value=float(_eval_node(op, [a.value])))
    if op == "NEG" and a.op == "NEG":
        return a.args[0]
    return Expr(op, (a,))


def _dispatch(op, npfn, *args):
    if any(isinstance(a, Expr) for a in args):
        return (_un if len(args) == 1 else _bin)(op, *args)
    return npfn(*args)


def _julia_max(a, b):
    return np.where(np.asarray(b) < np.asarray(a), a, b) if (np.ndim(a) or np.ndim(b)) else (a if b < a else b)


def _julia_min(a, b):
    return np.where(np.asarray(b) < np.asarray(a), b, a) if (np.ndim(a) or np.ndim(b)) else (b if b < a else a)


def exp(a): return _dispatch("EXP", np.exp, a)
def log(a): return _dispatch("LOG", np.log, a)
def sin(a): return _dispatch("SIN", np.sin, a)
def cos(a): return _dispatch("COS", np.cos, a)
def sqrt(a): return _dispatch("SQRT", np.sqrt, a)
def fabs(a): return _dispatch("ABS", np.abs, a)
def max(a, b): return _dispatch("MAX", _julia_max, a, b)  # noqa: A001 - mirrors Julia's max
def min(a, b): return _dispatch("MIN", _julia_min, a, b)  # noqa: A001


def ifelse(c, a, b):
    """``c ? a : b`` with ``c`` an indicator (nonzero = true)."""
    if any(isinstance(v, Expr) for v in (c, a, b)):
        return Expr("SELECT", (Expr.wrap(c), Expr.wrap(a), Expr.wrap(b)))
    return np.where(np.asarray(c) != 0, a, b)


# -- symbolic arguments ----------------------------------------------------------------
class _Indexed:
    def __init__(self, kind, n):
        self.kind, self.n = kind, n

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[k] for k in range(*i.indices(self.n))]
        i = int(i)
        if i < 0:
            i += self.n
        if not 0 <= i < self.n:
            raise IndexError(f"{'species' if self.kind == 'SPECIES' else 'parameter'} index {i} out of range 0..{self.n - 1}")
        return Expr(self.kind, value=i)

    def __len__(self):
        return self.n

    def __iter__(self):
        return (self[i] for i in range(self.n))


def symbols(ns, np_):
    """Symbolic ``(t, x, p)`` for tracing a callable with ``ns`` species and ``np_`` parameters."""
    return Expr("TIME"), _Indexed("SPECIES", ns), _Indexed("PARAM", np_)


def trace(fn, *args):
    out = fn(*args)
    return Expr.wrap(out)


# -- compile / evaluate ----------------------------------------------------------------
def compile_expr(e: Expr):
    """Postfix program as a list of ``(op, arg, c)`` tuples."""
    prog = []

    def emit(n):
        if n.op == "CONST":
            prog.append((OP["CONST"], 0, float(n.value)))
        elif n.op in ("PARAM", "SPECIES"):
            prog.append((OP[n.op], int(n.value), 0.0))
        elif n.op == "TIME":
            prog.append((OP["TIME"], 0, 0.0))
        else:
            for a in n.args:
                emit(a)
            prog.append((OP[n.op], 0, 0.0))

    emit(e)
    depth = peak = 0
    for op, _, _ in prog:
        if op in (OP["CONST"], OP["PARAM"], OP["SPECIES"], OP["TIME"]):
            depth += 1
        elif op == OP["SELECT"]:
            depth -= 2
        elif op in (OP[k] for k in _BINARY):
            depth -= 1
        peak = builtins_max(peak, depth)
    if peak > STACK_LIMIT:
        raise ValueError(f"propensity expression needs stack depth {peak} > {STACK_LIMIT}")
    return prog


def builtins_max(a, b):
    return a if a >= b else b


def _eval_node(op, v):
    if op == "ADD": return v[0] + v[1]
    if op == "SUB": return v[0] - v[1]
    if op == "MUL": return v[0] * v[1]
    if op == "DIV": return v[0] / v[1]
    if op == "POW": return np.power(v[0], v[1])
    if op == "NEG": return -v[0]
    if op == "EXP": return np.exp(v[0])
    if op == "LOG": return np.log(v[0])
    if op == "SIN": return np.sin(v[0])
    if op == "COS": return np.cos(v[0])
    if op == "SQRT": return np.sqrt(v[0])
    if op == "ABS": return np.abs(v[0])
    if op == "MAX": return _julia_max(v[0], v[1])
    if op == "MIN": return _julia_min(v[0], v[1])
    if op == "LT": return (np.asarray(v[0]) < v[1]) * 1.0
    if op == "LE": return (np.asarray(v[0]) <= v[1]) * 1.0
    if op == "GT": return (np.asarray(v[0]) > v[1]) * 1.0
    if op == "GE": return (np.asarray(v[0]) >= v[1]) * 1.0
    if op == "SELECT": return np.where(np.asarray(v[0]) != 0, v[1], v[2])
    raise ValueError(op)


def evaluate(e: Expr, t, x, p):
    """Evaluate a tree in numpy (``x``: sequence of per-species arrays or numbers)."""
    if e.op == "CONST":
        return e.value
    if e.op == "PARAM":
        return p[e.value]
    if e.op == "SPECIES":
        return np.asarray(x[e.value], dtype=np.float64)
    if e.op == "TIME":
        return t
    return _eval_node(e.op, [evaluate(a, t, x, p) for a in e.args])


def _factors(e: Expr, power, out):
    """Multiplicative factors of ``e`` as ``(factor, +1 | -1)`` pairs (-1: a divisor)."""
    if e.op == "MUL":
        _factors(e.args[0], power, out)
        _factors(e.args[1], power, out)
    elif e.op == "DIV":
        _factors(e.args[0], power, out)
        _factors(e.args[1], -power, out)
    elif e.op == "NEG":
        out.append((Expr("CONST", value=-1.0), 1))
        _factors(e.args[0], power, out)
    else:
        out.append((e, power))
    return out


def _product(factors):
    num = [f for f, k in factors if k > 0]
    den = [f for f, k in factors if k < 0]
    out = ONE
    for f in num:
        out = _bin("MUL", out, f)
    if den:
        d = den[0]
        for f in den[1:]:
            d = _bin("MUL", d, f)
        out = _bin("DIV", out, d)
    return out


def factor_time(e: Expr):
    """Split a joint law ``f(t, x, p)`` into ``tfactor(t, p) * statefactor(x, p)``.

    Catalyst hands every time-varying rate law over as a joint propensity
    (``src/cmemodel/catalyst_interface.jl:21-27``), and a joint propensity makes the
    generator re-evaluate every state's propensity whenever ``t`` changes
    (``src/fspmatrix/sparse/fspsparsematrix.jl:154-166``). Most laws are products such as
    ``k(t) * X`` or ``X / (1 + t)``; for those the matrix keeps the state factor and scales
    it by the time factor at each launch, as for a ``SeparableTimeVaryingPropensity``.

    ``e``'s product/quotient factors are grouped into those that read time (and no species)
    and the rest. Returns ``("ti", f)`` when ``e`` does not read time, ``("sep", tfactor,
    statefactor)`` when it factors, and ``None`` when a factor reads both time and the state
    (e.g. ``x / (1 + t x)``). The factored product equals ``e`` up to the rounding of a
    reordered product.
    """
    if not uses(e, "TIME"):
        return ("ti", e)
    tf, sf = [], []
    for f, k in _factors(e, 1, []):
        ut, us = uses(f, "TIME"), uses(f, "SPECIES")
        if ut and us:
            return None
        (tf if ut else sf).append((f, k))
    return ("sep", _product(tf), _product(sf))


def params_used(e: Expr, acc=None):
    acc = set() if acc is None else acc
    if e.op == "PARAM":
        acc.add(e.value)
    for a in e.args:
        params_used(a, acc)
    return acc


def uses(e: Expr, kind):
    return e.op == kind or any(uses(a, kind) for a in e.args)


ZERO = Expr("CONST", value=0.0)
ONE = Expr("CONST", value=1.0)


def diff(e: Expr, k: int) -> Expr:
    """Closed-form derivative of ``e`` with respect to parameter ``k``."""
    op = e.op
    if op == "CONST" or op == "SPECIES" or op == "TIME":
        return ZERO
    if op == "PARAM":
        return ONE if e.value == k else ZERO
    if k not in params_used(e):
        return ZERO
    a = e.args
    if op == "ADD":
        return _bin("ADD", diff(a[0], k), diff(a[1], k))
    if op == "SUB":
        return _bin("SUB", diff(a[0], k), diff(a[1], k))
    if op == "MUL":
        da, db = diff(a[0], k), diff(a[1], k)
        t1 = ZERO if _is_const(da, 0.0) else _bin("MUL", da, a[1])
        t2 = ZERO if _is_const(db, 0.0) else _bin("MUL", a[0], db)
        return _bin("ADD", t1, t2)
    if op == "DIV":
        da, db = diff(a[0], k), diff(a[1], k)
        t1 = ZERO if _is_const(da, 0.0) else _bin("DIV", da, a[1])
        if _is_const(db, 0.0):
            return t1
        t2 = _bin("DIV", _bin("MUL", a[0], db), _bin("MUL", a[1], a[1]))
        return _bin("SUB", t1, t2)
    if op == "POW":
        u, v = a
        du, dv = diff(u, k), diff(v, k)
        out = ZERO
        if not _is_const(du, 0.0):
            # v * u^(v-1) * du
            out = _bin("MUL", _bin("MUL", v, _bin("POW", u, _bin("SUB", v, ONE))), du)
        if not _is_const(dv, 0.0):
            # u^v log(u) dv, taken as 0 at u = 0 like ForwardDiff's power rule
            term = _bin("MUL", _bin("MUL", e, _un("LOG", u)), dv)
            out = _bin("ADD", out, Expr("SELECT", (_bin("GT", u, 0.0), term, ZERO)))
        return out
    if op == "NEG":
        return _un("NEG", diff(a[0], k))
    if op == "EXP":
        return _bin("MUL", e, diff(a[0], k))
    if op == "LOG":
        return _bin("DIV", diff(a[0], k), a[0])
    if op == "SIN":
        return _bin("MUL", _un("COS", a[0]), diff(a[0], k))
    if op == "COS":
        return _un("NEG", _bin("MUL", _un("SIN", a[0]), diff(a[0], k)))
    if op == "SQRT":
        return _bin("DIV", diff(a[0], k), _bin("MUL", Expr.wrap(2.0), e))
    if op == "ABS":
        return Expr("SELECT", (_bin("LT", a[0], 0.0), _un("NEG", diff(a[0], k)), diff(a[0], k)))
    if op == "MAX":
        # Julia: max(a, b) = b < a ? a : b
        return Expr("SELECT", (_bin("LT", a[1], a[0]), diff(a[0], k), diff(a[1], k)))
    if op == "MIN":
        return Expr("SELECT", (_bin("LT", a[1], a[0]), diff(a[1], k), diff(a[0], k)))
    if op in ("LT", "LE", "GT", "GE"):
        return ZERO
    if op == "SELECT":
        return Expr("SELECT", (a[0], diff(a[1], k), diff(a[2], k)))
    raise ValueError(op)
