"""ctypes binding of ``libnumcme_hip.so`` (the C ABI declared in ``include/numcme_hip.h``).

The library is loaded from this package directory. There is no fallback: if the shared
object is missing, or no HIP device is present, calls raise ``NumCmeHipError``.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libnumcme_hip.so"

NC_OK = 0
NC_ERR_ARG = 1
NC_ERR_HIP = 2
NC_ERR_ALLOC = 3
NC_ERR_OVERFLOW = 4
NC_ERR_UNSUPPORTED = 5
NC_ERR_NUMERIC = 6
NC_ERR_INTERNAL = 7
NC_ERR_NO_DEVICE = 8

NC_FP64 = 0
NC_FP32 = 1

NC_PROP_TIME_INVARIANT = 0
NC_PROP_SEPARABLE = 1
NC_PROP_JOINT = 2

NC_METHOD_AUTO = 0
NC_METHOD_KRYLOV = 1
NC_METHOD_RK45 = 2

NC_ADAPTER_RSTEP = 0
NC_ADAPTER_SELECTIVE_RSTEP = 1
NC_ADAPTER_LEARNED = 2


class NumCmeHipError(RuntimeError):
    """A failed library call; ``code`` is the C status."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class ArgumentError(NumCmeHipError, ValueError):
    """Invalid argument — the reference throws Julia's ``ArgumentError`` in these cases."""


class nc_instr(C.Structure):
    _fields_ = [("op", C.c_int32), ("arg", C.c_int32), ("c", C.c_double)]


class nc_matrix_info_t(C.Structure):
    _fields_ = [
        ("n_rows", C.c_int64),
        ("n_states", C.c_int64),
        ("n_sinks", C.c_int32),
        ("ell_width", C.c_int32),
        ("n_tv", C.c_int32),
        ("precision", C.c_int32),
        ("nnz_reference", C.c_int64),
        ("n_sink_entries", C.c_int64),
        ("bytes_per_matvec", C.c_double),
        ("csr_bytes_per_matvec", C.c_double),
        ("norm1", C.c_double),
        ("rows_per_thread", C.c_int32),
        ("computed_col_slots", C.c_int32),
        ("matrix_free", C.c_int32),
        ("matrix_free_available", C.c_int32),
    ]


class nc_solve_opts(C.Structure):
    _fields_ = [
        ("method", C.c_int32),
        ("krylov_m", C.c_int32),
        ("krylov_tol", C.c_double),
        ("rtol", C.c_double),
        ("atol", C.c_double),
        ("h0", C.c_double),
        ("max_steps", C.c_int64),
        ("time_kernels", C.c_int32),
        ("no_graphs", C.c_int32),
    ]


class nc_solve_stats(C.Structure):
    _fields_ = [
        ("steps", C.c_int64),
        ("rejected", C.c_int64),
        ("matvecs", C.c_int64),
        ("host_syncs", C.c_int64),
        ("matvecs_timed", C.c_int64),
        ("matvec_ms", C.c_double),
        ("error_estimate", C.c_double),
        ("adaptations", C.c_int64),
        ("learned_adaptations", C.c_int64),
        ("fallback_adaptations", C.c_int64),
        ("max_states", C.c_int64),
    ]


class nc_adapter(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("initial_step_count", C.c_int32),
        ("max_step_count", C.c_int32),
        ("dropstates", C.c_int32),
        ("mlp", C.c_void_p),
        ("threshold", C.c_double),
        ("fallback_dt", C.c_double),
    ]


class nc_adaptive_opts(C.Structure):
    _fields_ = [
        ("adapter", nc_adapter),
        ("precision", C.c_int32),
        ("max_adaptations", C.c_int32),
        ("fsptol", C.c_double),
        ("ode", nc_solve_opts),
        ("verbose", C.c_int32),
        ("records", C.c_void_p),
    ]


P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
D = C.c_double
PI32 = C.POINTER(C.c_int32)
PI64 = C.POINTER(C.c_int64)
PD = C.POINTER(C.c_double)
PU8 = C.POINTER(C.c_uint8)

# name -> argtypes (all functions return int status)
SIGNATURES = {
    "nc_last_error": [C.c_char_p, C.c_size_t],
    "nc_memory_trim": [],
    "nc_version": [PI32, PI32],
    "nc_device_name": [C.c_char_p, C.c_size_t],
    "nc_space_create": [I32, I32, PI32, I64, PI32, C.POINTER(P)],
    "nc_space_destroy": [P],
    "nc_space_expand": [P, I32, I32, PI32],
    "nc_space_delete": [P, I64, PI64],
    "nc_space_count": [P, PI64, PI32],
    "nc_space_get_states": [P, PI32],
    "nc_space_get_state_conn": [P, PI32],
    "nc_space_get_sink_conn": [P, PI32],
    "nc_space_lookup": [P, I64, PI32, PI64],
    "nc_model_create": [I32, I32, PI32, I32, PD, PI32, PI64, C.POINTER(nc_instr), C.POINTER(P)],
    "nc_model_destroy": [P],
    "nc_model_set_gradients": [P, PU8, PI64, C.POINTER(nc_instr)],
    "nc_model_eval": [P, I32, I32, D, I64, PI32, PD],
    "nc_matrix_create": [P, P, I32, C.POINTER(P)],
    "nc_matrix_destroy": [P],
    "nc_matrix_info": [P, C.POINTER(nc_matrix_info_t)],
    "nc_matrix_matvec": [P, D, PD, PD, I32],
    "nc_matrix_matvec_noalloc": [P, D, PD, PD, I32],
    "nc_matrix_export": [P, D, PI64, PI64, PI64, PD],
    "nc_matrix_bench_matvec": [P, D, I32, PD],
    "nc_matrix_bench_projection": [P, I32, I32, PD],
    "nc_matrix_set_rows_per_thread": [P, I32],
    "nc_matrix_set_matrix_free": [P, I32],
    "nc_matrix_autotune": [P, I32, PD],
    "nc_solve_opts_default": [C.POINTER(nc_solve_opts)],
    "nc_solver_create": [P, C.POINTER(nc_solve_opts), C.POINTER(P)],
    "nc_solver_destroy": [P],
    "nc_solver_set_state": [P, D, PD],
    "nc_solver_restart": [P],
    "nc_solver_run": [P, I32, PD],
    "nc_solver_get_snapshot": [P, I32, PD],
    "nc_solver_get_state": [P, PD, PD],
    "nc_solver_stats": [P, C.POINTER(nc_solve_stats)],
    "nc_solve_fixed": [P, C.POINTER(nc_solve_opts), PD, I32, PD, PD, C.POINTER(nc_solve_stats)],
    "nc_sensmatrix_create": [P, P, I32, C.POINTER(P)],
    "nc_sensmatrix_destroy": [P],
    "nc_sensmatrix_info": [P, PI64, PI32, PI32, PD],
    "nc_sensmatrix_matvec": [P, D, PD, PD],
    "nc_sensmatrix_bench_matvec": [P, D, I32, PD],
    "nc_solver_create_sens": [P, C.POINTER(nc_solve_opts), C.POINTER(P)],
    "nc_solve_sens_fixed": [P, C.POINTER(nc_solve_opts), PD, I32, PD, PD, C.POINTER(nc_solve_stats)],
    "nc_batch_create": [I32, C.POINTER(P), C.POINTER(P)],
    "nc_batch_create_models": [I32, C.POINTER(P), C.POINTER(C.c_void_p), PI64, I32, I32, C.POINTER(P)],
    "nc_batch_destroy": [P],
    "nc_batch_info": [P, PI64, PI64],
    "nc_batch_solve": [P, C.POINTER(nc_solve_opts), PD, I32, PD, C.POINTER(nc_solve_stats)],
    "nc_batch_get_snapshot": [P, I32, I32, PD],
    "nc_batch_get_snapshots": [P, I32, PD],
    "nc_mlp_create": [C.c_uint64, C.POINTER(P)],
    "nc_mlp_destroy": [P],
    "nc_mlp_get_params": [P, C.POINTER(C.c_float)],
    "nc_mlp_set_params": [P, C.POINTER(C.c_float)],
    "nc_mlp_forward": [P, I64, C.POINTER(C.c_float), C.POINTER(C.c_float)],
    "nc_mlp_grad": [P, I64, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), PD],
    "nc_mlp_train": [P, I64, C.POINTER(C.c_float), C.POINTER(C.c_float), I32, I32, C.c_float, C.c_uint64, PD],
    "nc_batch_autotune": [P, I32, C.POINTER(I32), PD],
    "nc_batch_set_rows_per_thread": [P, I32],
    "nc_mlp_bench": [P, I64, I32, I32, PD],
    "nc_adapter_features": [P, P, I64, PD, PD, C.c_double, C.c_double, C.POINTER(I64), C.POINTER(I64),
                            C.POINTER(C.c_float)],
    "nc_records_create": [I64, I64, C.c_double, C.POINTER(P)],
    "nc_records_destroy": [P],
    "nc_records_count": [P, C.POINTER(I64)],
    "nc_records_get": [P, C.POINTER(C.c_float), C.POINTER(C.c_float)],
    "nc_records_clear": [P],
    "nc_adapter_init": [P, C.POINTER(nc_adapter)],
    "nc_adapter_adapt": [P, C.POINTER(nc_adapter), I64, PD, PD, D, D, D, PI64, PI64],
    "nc_adaptive_opts_default": [C.POINTER(nc_adaptive_opts)],
    "nc_solve_adaptive": [P, I64, PI32, PD, D, D, I32, PD, C.POINTER(nc_adaptive_opts), C.POINTER(P),
                          C.POINTER(nc_solve_stats)],
    "nc_solve_adaptive_batch": [I32, P, PI64, PI32, PD, D, D, I32, PD, C.POINTER(nc_adaptive_opts), I32, P, P],
    "nc_solve_sens_adaptive": [P, I64, PI32, PD, PD, D, D, I32, PD, C.POINTER(nc_adaptive_opts), C.POINTER(P),
                               C.POINTER(nc_solve_stats)],
    "nc_output_get_sens": [P, I32, I32, PD, PD],
    "nc_output_destroy": [P],
    "nc_output_count": [P, PI32, PI32],
    "nc_output_info": [P, I32, PD, PI64, PI32],
    "nc_output_get": [P, I32, PI32, PD, PD],
    "nc_expm_dense": [I32, PD, PD],
    "nc_fspvec_create": [I32, I64, PI32, PD, C.POINTER(P)],
    "nc_fspvec_destroy": [P],
    "nc_fspvec_size": [P, PI64, PI32],
    "nc_fspvec_get": [P, PI32, PD],
    "nc_fspvec_sum": [P, PD],
    "nc_fspvec_marginal": [P, I32, PI32, C.POINTER(P)],
    "nc_fspvec_dense": [P, PI64, I64, PD],
    "nc_solver_get_fspvec": [P, I32, I32, P, C.POINTER(P), PD],
    "nc_output_get_fspvec": [P, I32, I32, C.POINTER(P), PD],
}

_lib = None


def lib():
    """Load the shared library once (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise NumCmeHipError(NC_ERR_INTERNAL, f"{LIB_PATH} is missing; build it with numcme.jl_amd/build.py")
        L = C.CDLL(str(LIB_PATH))
        for name, argtypes in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = argtypes
            fn.restype = C.c_int
        _lib = L
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(4096)
    lib().nc_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc: int):
    if rc != NC_OK:
        msg = last_error()
        if rc == NC_ERR_ARG:
            raise ArgumentError(rc, msg)
        raise NumCmeHipError(rc, msg)


def call(name, *args):
    check(getattr(lib(), name)(*args))


def ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def device_name() -> str:
    buf = C.create_string_buffer(256)
    call("nc_device_name", buf, len(buf))
    return buf.value.decode()


def expm_dense(a) -> np.ndarray:
    """Host dense matrix exponential used by the Krylov integrator (no GPU needed)."""
    a = f64(a)
    n = a.shape[0]
    out = np.empty_like(a)
    call("nc_expm_dense", n, ptr(a, C.c_double), ptr(out, C.c_double))
    return out


class Handle:
    """Owns one opaque library handle and destroys it when collected."""

    _destroy = None

    def __init__(self, h):
        self._h = h

    @property
    def handle(self):
        if self._h is None:
            raise NumCmeHipError(NC_ERR_ARG, f"{type(self).__name__} has been freed")
        return self._h

    def free(self):
        if getattr(self, "_h", None) is not None and self._destroy:
            try:
                getattr(lib(), self._destroy)(self._h)
            finally:
                self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
