"""Timed CPU baseline driver (TEST INFRASTRUCTURE: bench.py's cpu_baseline leg only).

Builds ``cpu_ref.c`` with gcc into ``oracle/_build/libcpu_ref.so`` and runs the same
Krylov expv algorithm as the GPU integrator (Expokit step control, m = 30) on the
CPU with the reference's CSC column-scatter mat-vec, single-threaded. The step
control and the small Hessenberg exponential run in Python/scipy (negligible cost);
the timed work is the C mat-vec and Arnoldi.
"""
from __future__ import annotations

import ctypes as C
import subprocess
import time
from pathlib import Path

import numpy as np
import scipy.linalg as sla

HERE = Path(__file__).resolve().parent
SO = HERE / "_build" / "libcpu_ref.so"


def build():
    (HERE / "_build").mkdir(exist_ok=True)
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fPIC", "-shared", str(HERE / "cpu_ref.c"), "-o", str(SO), "-lm"]
    subprocess.run(cmd, check=True, capture_output=True)
    return SO


def _lib():
    if not SO.exists():
        build()
    L = C.CDLL(str(SO))
    P = C.c_void_p
    L.csc_matvec.argtypes = [C.c_int64, P, P, P, P, P]
    L.arnoldi_cgs2.argtypes = [C.c_int64, P, P, P, C.c_int, P, C.c_double, P, P, P, C.POINTER(C.c_double)]
    L.arnoldi_cgs2.restype = C.c_int
    L.combine.argtypes = [C.c_int64, C.c_int, P, P, P]
    L.combine.restype = C.c_double
    return L


def _round_step(x):
    s = 10.0 ** (np.floor(np.log10(x)) - 1)
    return np.ceil(x / s) * s


class CscExpv:
    """Krylov expv on a scipy CSC matrix with the C kernels."""

    def __init__(self, A, m=30, tol=1e-12):
        A = A.tocsc()
        A.sort_indices()
        self.n = A.shape[0]
        self.colptr = np.ascontiguousarray(A.indptr, dtype=np.int64)
        self.rowval = np.ascontiguousarray(A.indices, dtype=np.int32)
        self.nzval = np.ascontiguousarray(A.data, dtype=np.float64)
        self.m, self.tol = m, tol
        self.anorm = 2 * np.abs(A.diagonal()).max()
        self.L = _lib()
        self.V = np.zeros((m + 1) * self.n)
        self.H = np.zeros((m + 2) * (m + 2))
        self.p = np.zeros(self.n)

    def matvec(self, v, out):
        a = lambda x: x.ctypes.data_as(C.c_void_p)
        self.L.csc_matvec(self.n, a(self.colptr), a(self.rowval), a(self.nzval), a(v), a(out))

    def run(self, w, t_total, max_steps=None):
        """Advance w by t_total; returns (w, steps, seconds) — stops early after max_steps."""
        a = lambda x: x.ctypes.data_as(C.c_void_p)
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
            mb = self.L.arnoldi_cgs2(n, a(self.colptr), a(self.rowval), a(self.nzval), m, a(w), beta, a(self.V),
                                     a(self.H), a(self.p), C.byref(av))
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
            beta = self.L.combine(n, mx, a(self.V), a(c), a(w))
            t_now += t_step
            steps += 1
            if k1:
                t_new = _round_step(0.9 * t_step * (t_step * tol / err) ** xm)
        return w, steps, time.perf_counter() - t0
