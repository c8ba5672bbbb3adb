"""C restatement driver (TEST INFRASTRUCTURE: bench.py's cpu_baseline leg and the parity
tests' large exact solutions only).

Builds ``cpu_ref.c`` with gcc (OpenMP) into ``oracle/_build/libcpu_ref.so``.

* ``Expv``: the Krylov expv algorithm of the GPU integrator (Expokit step control, m = 30)
  with the C kernels: either the reference's CSC column-scatter mat-vec on one thread
  (``layout="csc"``), or a row-parallel CSR mat-vec with the vector work over
  ``threads`` OpenMP threads (``layout="csr"``). The step control and the small Hessenberg
  exponential run in Python/scipy (negligible cost).
* ``uniformization``: the uniformization series of ``oracle.solve.uniformization`` with the
  C CSR mat-vec over all threads, for exact solutions of large generators.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
import time
from pathlib import Path

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sps

HERE = Path(__file__).resolve().parent
SO = HERE / "_build" / "libcpu_ref.so"


def build():
    (HERE / "_build").mkdir(exist_ok=True)
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", str(HERE / "cpu_ref.c"), "-o", str(SO),
           "-lm"]
    subprocess.run(cmd, check=True, capture_output=True)
    return SO


_L = None


def _lib():
    global _L
    if _L is None:
        if not SO.exists():
            build()
        L = C.CDLL(str(SO))
        P = C.c_void_p
        L.cpu_threads.restype = C.c_int
        L.csc_matvec.argtypes = [C.c_int64, P, P, P, P, P]
        L.csr_matvec.argtypes = [C.c_int64, P, P, P, C.c_int, P, P]
        L.krylov_arnoldi.argtypes = [C.c_int64, P, P, P, C.c_int, C.c_int, C.c_int, P, C.c_double, P, P, P,
                                     C.POINTER(C.c_double)]
        L.krylov_arnoldi.restype = C.c_int
        L.krylov_combine.argtypes = [C.c_int64, C.c_int, C.c_int, P, P, P]
        L.krylov_combine.restype = C.c_double
        L.uniformize.argtypes = [C.c_int64, P, P, P, C.c_int, P, C.c_int, P, P, P, P]
        _L = L
    return _L


def cpu_count():
    """Cores this process may run on (the affinity mask; the OpenMP default team size)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _a(x):
    return x.ctypes.data_as(C.c_void_p)


def _round_step(x):
    s = 10.0 ** (np.floor(np.log10(x)) - 1)
    return np.ceil(x / s) * s


class Expv:
    """Krylov expv on a scipy sparse generator with the C kernels."""

    def __init__(self, A, m=30, tol=1e-12, layout="csc", threads=1):
        if layout == "csc":
            M = sps.csc_matrix(A)
            threads = 1
        else:
            M = sps.csr_matrix(A)
        M.sort_indices()
        self.n = M.shape[0]
        self.csr = 0 if layout == "csc" else 1
        self.threads = int(threads)
        self.ptr = np.ascontiguousarray(M.indptr, dtype=np.int64)
        self.idx = np.ascontiguousarray(M.indices, dtype=np.int32)
        self.val = np.ascontiguousarray(M.data, dtype=np.float64)
        self.m, self.tol = m, tol
        self.anorm = 2 * np.abs(M.diagonal()).max()
        self.L = _lib()
        self.V = np.zeros((m + 1) * self.n)
        self.H = np.zeros((m + 2) * (m + 2))
        self.p = np.zeros(self.n)

    def matvec(self, v, out):
        if self.csr:
            self.L.csr_matvec(self.n, _a(self.ptr), _a(self.idx), _a(self.val), self.threads, _a(v), _a(out))
        else:
            self.L.csc_matvec(self.n, _a(self.ptr), _a(self.idx), _a(self.val), _a(v), _a(out))

    def run(self, w, t_total, max_steps=None):
        """Advance w by t_total; returns (w, steps, seconds) — stops early after max_steps."""
        m, tol, n = self.m, self.tol, self.n
        w = np.array(w, dtype=np.float64)
        beta = np.linalg.norm(w)
        fact = ((m + 1) / np.e) ** (m + 1) * np.sqrt(2 * np.pi * (m + 1))
        t_new = _round_step((1 / self.anorm) * ((fact * tol) / (4 * beta * self.anorm)) ** (1 / m))
        t_now, steps = 0.0, 0
        t0 = time.perf_counter()
        while t_now < t_total * (1 - 1e-13):
            if max_steps is not None and steps >= max_steps:
                break
            t_step = min(t_total - t_now, t_new)
            av = C.c_double()
            mb = self.L.krylov_arnoldi(n, _a(self.ptr), _a(self.idx), _a(self.val), self.csr, self.threads, m, _a(w),
                                       beta, _a(self.V), _a(self.H), _a(self.p), C.byref(av))
            H = self.H.reshape(m + 2, m + 2)
            k1 = 2 if mb == m else 0
            if k1:
                H[m + 1, m] = 1.0
            else:
                t_step = t_total - t_now
            xm = 1.0 / m
            while True:
                mx = mb + k1
                F = sla.expm(t_step * H[:mx, :mx])
                if k1 == 0:
                    err = 1e-7
                    break
                phi1, phi2 = abs(beta * F[m, 0]), abs(beta * F[m + 1, 0] * av.value)
                if phi1 > 10 * phi2:
                    err, xm = phi2, 1.0 / m
                elif phi1 > phi2:
                    err, xm = phi1 * phi2 / (phi1 - phi2), 1.0 / m
                else:
                    err, xm = phi1, 1.0 / (m - 1)
                if err <= 1.2 * t_step * tol:
                    break
                t_step = _round_step(0.9 * t_step * (t_step * tol / err) ** xm)
            mx = mb + max(0, k1 - 1)
            c = np.ascontiguousarray(beta * F[:mx, 0])
            beta = self.L.krylov_combine(n, self.threads, mx, _a(self.V), _a(c), _a(w))
            t_now += t_step
            steps += 1
            if k1:
                t_new = _round_step(0.9 * t_step * (t_step * tol / err) ** xm)
        return w, steps, time.perf_counter() - t0

    def solve(self, u0, ts):
        """Snapshots at the ascending times ts (from ts[0]); returns (list, steps, seconds)."""
        out, u, steps, secs = [np.array(u0, dtype=np.float64)], np.array(u0, dtype=np.float64), 0, 0.0
        for a, b in zip(ts[:-1], ts[1:]):
            u, st, s = self.run(u, b - a)
            steps, secs = steps + st, secs + s
            out.append(u.copy())
        return out, steps, secs


def poisson_weights(lt, tail=1e-14):
    """e^{-lt} lt^k / k! for k = 0..kmax, kmax where the weights left out sum below tail."""
    from scipy.stats import poisson

    kmax = int(poisson.isf(tail, lt)) + 2
    mode = int(lt)
    w = np.empty(kmax + 1)
    w[min(mode, kmax)] = 1.0
    for k in range(mode + 1, kmax + 1):
        w[k] = w[k - 1] * lt / k
    for k in range(min(mode, kmax + 1) - 1, -1, -1):
        w[k] = w[k + 1] * (k + 1) / lt
    logw_mode = -lt + mode * math.log(lt) - math.lgamma(mode + 1)
    return w * math.exp(logw_mode) if mode <= kmax else w / w.sum()


def uniformization(A, u0, ts, threads=None, tail=1e-14):
    """oracle.solve.uniformization with the series summed by the C kernels."""
    A = sps.csr_matrix(A)
    L = float(np.abs(A.diagonal()).max()) * 1.0000001 or 1.0
    P = sps.csr_matrix(sps.identity(A.shape[0], format="csr") + A / L)
    P.sort_indices()
    ptr = np.ascontiguousarray(P.indptr, dtype=np.int64)
    idx = np.ascontiguousarray(P.indices, dtype=np.int32)
    val = np.ascontiguousarray(P.data, dtype=np.float64)
    n = A.shape[0]
    lib = _lib()
    th = int(threads or cpu_count())
    u = np.ascontiguousarray(u0, dtype=np.float64).copy()
    out = [u.copy()]
    acc, term, tmp = np.empty(n), np.empty(n), np.empty(n)
    for a, b in zip(ts[:-1], ts[1:]):
        lt = L * (b - a)
        if lt == 0.0:
            out.append(u.copy())
            continue
        w = np.ascontiguousarray(poisson_weights(lt, tail))
        lib.uniformize(n, _a(ptr), _a(idx), _a(val), th, _a(u), len(w) - 1, _a(w), _a(acc), _a(term), _a(tmp))
        u = acc.copy()
        out.append(u.copy())
    return out
