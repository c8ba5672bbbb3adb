"""Benchmark: FSP solves/s on one MI355X, with the mat-vec roofline and a CPU baseline.

Workload (one "step" = one complete solve, from model upload to snapshot download): the
two-species toggle switch on the fixed state space 0..1023 x 0..1023 (1,048,576 states +
4 sinks, 5,240,832 reference nonzeros), production 50/(1+(S/20)^2.5)+1, degradation 1 per
molecule, initial state (0,0), outputs at t = 0,1,...,10, device Krylov expv with m = 30.
Each step traces and uploads the model, builds the state space, the generator and the
solver on the GPU, integrates, and downloads the 11 snapshots. Storage FP32 (matrix values
and Krylov basis) with FP64 diagonal, accumulation and master state; the l1 distance of
every snapshot to the FP64 GPU solve at tolerance 1e-12 is measured in the same run (the
FP64 path itself is checked against the exact oracle by the test suite).

    python bench.py [--steps K] [--warmup W] [--precision fp32|fp64] [--no-cpu]

Prints one JSON line. Besides the headline, the line carries the other quantities the
metric names: the batched form against the plain FP64 single-solve path (C4), the learned
adapter against the plain FP64 adapter (C5), the SpMV roofline, and the CPU restatement.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "numcme.jl_amd"))

HBM_PEAK_GBPS = 8000.0
TRAFFIC_PROFILE = ROOT / "profiles" / "k_matvec_toggle1024_fp32.md"
PROJECTION_PROFILE = ROOT / "profiles" / "k_matvec_projection_toggle1024_fp32.md"


def measured_traffic(rows_per_thread, precision, matrix_free, projection=False):
    """HBM bytes per mat-vec launch from the committed rocprofv3 counter runs (FETCH_SIZE
    plus WRITE_SIZE, calibrated on a copy kernel), for the mode and launch shape in use
    (projection: the mat-vec with the Krylov projection epilogue, averaged over a solve's
    launches); None when no counter run covers it."""
    prof = PROJECTION_PROFILE if projection else TRAFFIC_PROFILE
    if precision != "fp32" or not prof.exists():
        return None
    mode = "projection" if projection else {0: "stored", 1: "computed_values", 2: "matrix_free"}[int(matrix_free)]
    key = f"traffic_bytes_per_launch mode={mode} rows_per_thread={rows_per_thread}:"
    for line in prof.read_text().splitlines():
        if line.startswith(key):
            return float(line[len(key):].strip())
    return None


def toggle_model(nh):
    S = np.array([[1, -1, 0, 0], [0, 0, 1, -1]])
    props = [
        nh.propensity(lambda x, p: p[3] + p[0] / (1.0 + (x[1] / p[1]) ** p[2])),
        nh.propensity(lambda x, p: p[4] * x[0]),
        nh.propensity(lambda x, p: p[3] + p[0] / (1.0 + (x[0] / p[1]) ** p[2])),
        nh.propensity(lambda x, p: p[4] * x[1]),
    ]
    return nh.CmeModel(S, props, [50.0, 20.0, 2.5, 1.0, 1.0])


def box_states(n):
    """0..n-1 squared, species 0 fastest"""
    out = np.empty((n * n, 2), dtype=np.int64)
    out[:, 0] = np.tile(np.arange(n), n)
    out[:, 1] = np.repeat(np.arange(n), n)
    return out


class Solver:
    def __init__(self, L, A, precision, time_kernels, graphs=True, tol=1e-12):
        o = L.nc_solve_opts()
        L.call("nc_solve_opts_default", C.byref(o))
        o.method = L.NC_METHOD_KRYLOV
        o.krylov_m = 30
        o.krylov_tol = tol
        o.time_kernels = int(time_kernels)
        o.no_graphs = 0 if graphs else 1
        h = C.c_void_p()
        L.call("nc_solver_create", A.handle, C.byref(o), C.byref(h))
        self.L, self.h, self.A = L, h.value, A

    def set_state(self, u0):
        u0 = np.ascontiguousarray(u0, dtype=np.float64)
        self.L.call("nc_solver_set_state", self.h, 0.0, self.L.ptr(u0, C.c_double))

    def solve(self, ts):
        self.L.call("nc_solver_restart", self.h)
        self.L.call("nc_solver_run", self.h, len(ts), self.L.ptr(ts, C.c_double))

    def stats(self):
        st = self.L.nc_solve_stats()
        self.L.call("nc_solver_stats", self.h, C.byref(st))
        return {f: getattr(st, f) for f, _ in self.L.nc_solve_stats._fields_}

    def snapshot(self, k, n):
        out = np.empty(n)
        self.L.call("nc_solver_get_snapshot", self.h, k, self.L.ptr(out, C.c_double))
        return out

    def free(self):
        self.L.call("nc_solver_destroy", self.h)


def toggle_oracle_csc(theta, n=1024):
    """The oracle's FP64 toggle generator on the n x n box (bench-side CPU restatement input)."""
    from oracle.fspmatrix import FspMatrix, TimeInvariant
    from oracle.statespace import StateSpace

    S = np.array([[1, -1, 0, 0], [0, 0, 1, -1]])
    th = theta
    hill = lambda y: th[3] + th[0] / (1.0 + (y / th[1]) ** th[2])
    props = [TimeInvariant(lambda x, p: hill(x[1])), TimeInvariant(lambda x, p: th[4] * x[0]),
             TimeInvariant(lambda x, p: hill(x[0])), TimeInvariant(lambda x, p: th[4] * x[1])]
    return FspMatrix(StateSpace(S, box_states(n)), props, th).to_csc(0.0)


def cpu_baseline(model_theta, krylov_tol, gpu_fp64_steps, st_fraction=0.05, log=None):
    """The C restatement of the reference's CPU path (oracle/cpu_ref.c) on the host cores:
    the same Krylov expv algorithm, FP64, on C2 (the full solve with the row-parallel CSR
    mat-vec on every core; the first twentieth of the solve's steps with the reference's
    one-thread CSC scatter, scaled by the step count), on C1 (full solve, one thread) and on
    a sample of the C4 table (all cores)."""
    from oracle import cpu_ref
    from oracle.fspmatrix import FspMatrix
    from oracle.statespace import StateSpace

    say = log or (lambda m: None)
    cpu_ref.build()
    cores = cpu_ref.cpu_count()
    ts = np.arange(0.0, 11.0)
    A = toggle_oracle_csc(model_theta)
    u0 = np.zeros(A.shape[0])
    u0[0] = 1.0
    # all cores, full solve
    ex = cpu_ref.Expv(A, m=30, tol=krylov_tol, layout="csr", threads=cores)
    _, steps_mt, secs_mt = ex.solve(u0, ts)
    say(f"cpu: C2 full solve on {cores} threads {secs_mt:.1f}s ({steps_mt} steps)")
    # one thread, the reference's CSC scatter, a quarter of the steps (from t = 0)
    ex1 = cpu_ref.Expv(A, m=30, tol=krylov_tol, layout="csc")
    k1 = max(1, int(round(st_fraction * steps_mt)))
    _, st1, secs1 = ex1.run(u0, 10.0, max_steps=k1)
    per_step1 = secs1 / max(st1, 1)
    say(f"cpu: C2 one thread {st1} steps {secs1:.1f}s")
    # C1: telegraph on 800 states, one thread, full solve
    S1 = np.array([[-1, 1, 0, 0], [1, -1, 0, 0], [0, 0, 1, -1]])
    from oracle.fspmatrix import TimeInvariant
    th1 = (0.05, 0.1, 5.0, 0.05)
    props1 = [TimeInvariant(lambda x, p: p[0] * x[0]), TimeInvariant(lambda x, p: p[1] * x[1]),
              TimeInvariant(lambda x, p: p[2] * x[1]), TimeInvariant(lambda x, p: p[3] * x[2])]
    st_c1 = [[1, 0, m] for m in range(400)] + [[0, 1, m] for m in range(400)]
    A1 = FspMatrix(StateSpace(S1, st_c1), props1, np.asarray(th1)).to_csc(0.0)
    w1 = np.zeros(A1.shape[0])
    w1[0] = 1.0
    _, c1_steps, c1_secs = cpu_ref.Expv(A1, m=30, tol=krylov_tol, layout="csc").solve(w1, np.arange(0.0, 201.0, 10.0))
    # C4: 8 instances evenly spaced by size, all cores, full solves
    table = json.loads(C4_TABLE.read_text())
    order = np.argsort([t["states"] for t in table], kind="stable")
    sample = order[np.linspace(0, len(order) - 1, 8).round().astype(int)]
    c4_secs = 0.0
    for i in sample:
        Ai = c4_oracle_csc(table[i])
        ui = np.zeros(Ai.shape[0])
        ui[0] = 1.0
        _, _, sec = cpu_ref.Expv(Ai, m=30, tol=krylov_tol, layout="csr", threads=cores).solve(ui, ts)
        c4_secs += sec
    say(f"cpu: C4 sample {c4_secs:.1f}s")
    return {
        "value": 1.0 / secs_mt,
        "unit": "solves/s",
        "cores": cores,
        "kind": "port",
        "label": "restatement, not reference (Julia is not installed on this machine)",
        "sample": f"the full FP64 C2 solve ({steps_mt} Krylov steps, m=30, tol {krylov_tol:g}) with a row-parallel "
                  f"CSR mat-vec and vector work on {cores} OpenMP threads, {secs_mt:.1f} s",
        "one_thread_csc": {
            "solves_per_s": 1.0 / (per_step1 * steps_mt), "cores": 1,
            "sample": f"{st1} of the {steps_mt} Krylov steps ({st1 / steps_mt:.0%}) of the same FP64 solve with "
                      f"the reference's CSC column-scatter mat-vec, {secs1:.1f} s, scaled by the step count"},
        "c1_one_thread_solves_per_s": 1.0 / c1_secs,
        "c1_sample": f"the full C1 solve (800 states, t = 0:10:200, {c1_steps} Krylov steps), one thread, CSC",
        "c4_all_cores_solves_per_s": len(sample) / c4_secs,
        "c4_sample": f"{len(sample)} C4 instances evenly spaced by size, full FP64 solves, {cores} threads, "
                     f"{c4_secs:.1f} s",
        "gpu_fp64_krylov_steps": gpu_fp64_steps,
    }


def gene4_model(nh):
    """Four gene states (one-hot G0..G3), nuclear and cytoplasmic mRNA; k21 pulse-modulated:
    k21(t) = max(0, k21 - 0.12 s(t)), s(t) = (1 - e^{-0.006 t}) e^{-0.009 t}; 12 parameters
    (k01, k10, k12, k21, k23, k32, four transcription rates, export, cytoplasmic decay)."""
    S = np.zeros((6, 12), dtype=np.int64)
    for r, (a, b) in enumerate([(0, 1), (1, 0), (1, 2), (2, 1), (2, 3), (3, 2)]):
        S[a, r], S[b, r] = -1, 1
    S[4, 6:10] = 1
    S[4, 10], S[5, 10] = -1, 1
    S[5, 11] = -1
    nm = nh.nmath
    sig = lambda t: (1.0 - nm.exp(-6e-3 * t)) * nm.exp(-9e-3 * t)
    props = [nh.propensity(lambda x, p: p[0] * x[0]), nh.propensity(lambda x, p: p[1] * x[1]),
             nh.propensity(lambda x, p: p[2] * x[1]),
             nh.propensity(lambda x, p: x[2], tfactor=lambda t, p: nm.max(0.0, p[3] - 0.12 * sig(t))),
             nh.propensity(lambda x, p: p[4] * x[2]), nh.propensity(lambda x, p: p[5] * x[3]),
             nh.propensity(lambda x, p: p[6] * x[0]), nh.propensity(lambda x, p: p[7] * x[1]),
             nh.propensity(lambda x, p: p[8] * x[2]), nh.propensity(lambda x, p: p[9] * x[3]),
             nh.propensity(lambda x, p: p[10] * x[4]), nh.propensity(lambda x, p: p[11] * x[5])]
    theta = [1.3e-3, 3.2e-3, 5.9e-3, 2e-2, 1.5e-3, 4.6e-3, 0.0, 1e-2, 0.2, 1.5, 0.05, 1e-3]
    return nh.CmeModelWithSensitivity(nh.CmeModel(S, props, theta))


def gene4_states(nmax):
    g = np.repeat(np.eye(4, dtype=np.int64), (nmax + 1) ** 2, axis=0)
    c, n = np.meshgrid(np.arange(nmax + 1), np.arange(nmax + 1), indexing="ij")
    mrna = np.tile(np.stack([n.ravel(), c.ravel()], axis=1), (4, 1))
    return np.concatenate([g, mrna], axis=1)


def c3_sensitivity(nh, L):
    """Config C3: 12-parameter forward sensitivities on 262,144 states, t = 0:60:900, RK45 on
    the device (the generator is time-varying), FP32 storage against FP64."""
    model = gene4_model(nh)
    states = gene4_states(255)
    space = nh.StateSpaceSparse(model.get_stoich_matrix(), states)
    ts = np.arange(0.0, 901.0, 60.0)
    res = {"workload": "4 gene states x 256 nuclear x 256 cytoplasmic mRNA (262,144 states + 12 sinks), "
                       "12-parameter forward sensitivities (13 blocks), t = 0:60:900, device RK45 rtol 1e-8"}
    finals = {}
    for prec in ("fp64", "fp32"):
        SA = nh.ForwardSensFspMatrixSparse(model, space, precision=prec)
        n, P = SA.rowcount, SA.parameter_count
        u0 = np.zeros(n * (P + 1))
        u0[0] = 1.0  # delta at (G0, 0, 0): the first state of the list
        o = L.nc_solve_opts()
        L.call("nc_solve_opts_default", C.byref(o))
        o.method, o.rtol, o.atol = L.NC_METHOD_RK45, 1e-8, 1e-14
        h = C.c_void_p()
        L.call("nc_solver_create_sens", SA.handle, C.byref(o), C.byref(h))
        L.call("nc_solver_set_state", h, 0.0, L.ptr(u0, C.c_double))
        t0 = time.perf_counter()
        L.call("nc_solver_run", h, len(ts), L.ptr(ts, C.c_double))
        res[f"{prec}_solve_s"] = time.perf_counter() - t0
        st = L.nc_solve_stats()
        L.call("nc_solver_stats", h, C.byref(st))
        res[f"{prec}_rk_steps"] = st.steps
        res[f"{prec}_sens_matvecs"] = st.matvecs
        fin = np.empty(n * (P + 1))
        L.call("nc_solver_get_snapshot", h, len(ts) - 1, L.ptr(fin, C.c_double))
        finals[prec] = fin.reshape(P + 1, n)
        L.call("nc_solver_destroy", h)
        ms = SA.bench_matvec(reps=50, t=300.0)
        res[f"{prec}_sens_matvec_us"] = ms * 1e3
        res[f"{prec}_sens_bytes_per_matvec"] = SA.bytes_per_matvec
        res[f"{prec}_sens_matvec_GBps"] = SA.bytes_per_matvec / (ms * 1e-3) / 1e9
        SA.free()
    a, b = finals["fp64"], finals["fp32"]
    res["l1_p_fp32_vs_fp64"] = float(np.abs(a[0] - b[0]).sum())
    res["max_rel_S_fp32_vs_fp64"] = float(max(np.abs(a[k] - b[k]).max() / max(np.abs(a[k]).max(), 1e-300)
                                              for k in range(1, a.shape[0])))
    # d/dtheta of the total mass is zero: sum of each sensitivity block (states + sinks)
    res["max_abs_mass_derivative_fp64"] = float(np.abs(a[1:].sum(axis=1)).max())
    return res


C4_TABLE = ROOT / "tests" / "golden" / "c4_instances.json"


_C4_MODELS = {}


def c4_model_states(nh, spec):
    """The model (parameter values of the instance, networks traced once) and the state
    list of one batched instance (telegraph or toggle)."""
    th = spec["theta"]
    nmax = spec["nmax"]
    if spec["kind"] not in _C4_MODELS:
        if spec["kind"] == "telegraph":
            S = np.array([[-1, 1, 0, 0], [1, -1, 0, 0], [0, 0, 1, -1]])
            props = [nh.propensity(lambda x, p: p[0] * x[0]), nh.propensity(lambda x, p: p[1] * x[1]),
                     nh.propensity(lambda x, p: p[2] * x[1]), nh.propensity(lambda x, p: p[3] * x[2])]
            _C4_MODELS["telegraph"] = nh.CmeModel(S, props, th)
        else:
            _C4_MODELS["toggle"] = toggle_model(nh).with_parameters(th)
    model = _C4_MODELS[spec["kind"]].with_parameters(th)
    if spec["kind"] == "telegraph":
        m = np.arange(nmax + 1)
        one, zero = np.ones_like(m), np.zeros_like(m)
        states = np.concatenate([np.stack([one, zero, m], axis=1), np.stack([zero, one, m], axis=1)])
    else:
        states = box_states(nmax + 1)
    return model, states


def c4_instance(nh, spec, precision):
    """Space, generator and initial vector of one batched instance (telegraph or toggle).
    The two networks are traced once; each instance takes its parameter values
    (CmeModel.with_parameters) and uploads its own device model."""
    model, states = c4_model_states(nh, spec)
    sp = nh.StateSpaceSparse(model.stoich_matrix, states)
    A = nh.FspMatrixSparse(sp, model.propensities, model.parameters, precision=precision, model=model)
    u0 = np.zeros(len(states) + model.stoich_matrix.shape[1])
    u0[0] = 1.0  # telegraph (G_off, 0) / toggle (0, 0): the first state of the list
    return A, u0


def c4_oracle_csc(spec):
    """The oracle's FP64 generator of one C4 instance (the CPU restatement's input)."""
    from oracle.fspmatrix import FspMatrix, TimeInvariant
    from oracle.statespace import StateSpace

    th, nmax = np.asarray(spec["theta"], dtype=float), spec["nmax"]
    if spec["kind"] == "telegraph":
        S = np.array([[-1, 1, 0, 0], [1, -1, 0, 0], [0, 0, 1, -1]])
        props = [TimeInvariant(lambda x, p: p[0] * x[0]), TimeInvariant(lambda x, p: p[1] * x[1]),
                 TimeInvariant(lambda x, p: p[2] * x[1]), TimeInvariant(lambda x, p: p[3] * x[2])]
        states = [[1, 0, m] for m in range(nmax + 1)] + [[0, 1, m] for m in range(nmax + 1)]
    else:
        S = np.array([[1, -1, 0, 0], [0, 0, 1, -1]])
        hill = lambda y, p: p[3] + p[0] / (1.0 + (y / p[1]) ** p[2])
        props = [TimeInvariant(lambda x, p: hill(x[1], p)), TimeInvariant(lambda x, p: p[4] * x[0]),
                 TimeInvariant(lambda x, p: hill(x[0], p)), TimeInvariant(lambda x, p: p[4] * x[1])]
        states = box_states(nmax + 1)
    return FspMatrix(StateSpace(S, states), props, th).to_csc(0.0)


def c4_batch(nh, L, krylov_tol, graphs=True, log=None, plain_leg=True):
    """Config C4: the 512 uneven instances in one batched FP32 solve against the plain FP64
    single-solve path, both timed on one clock: from the upload of every instance's state
    space and generator to the download of every snapshot. plain_leg=False: the batched
    solves alone (profiler runs)."""
    say = log or (lambda m: None)
    table = json.loads(C4_TABLE.read_text())
    ts = np.arange(0.0, 11.0)

    def batched():
        # the models and state lists, then every space and generator built by the
        # library's worker threads straight into the batch (nc_batch_create_models)
        t0 = time.perf_counter()
        models, states, u0s = [], [], []
        for spec in table:
            model, st = c4_model_states(nh, spec)
            models.append(model)
            states.append(st)
            u0 = np.zeros(len(st) + model.stoich_matrix.shape[1])
            u0[0] = 1.0  # telegraph (G_off, 0) / toggle (0, 0): the first state of the list
            u0s.append(u0)
        B = nh.BatchFsp.from_models(models, states, "fp32")
        t1 = time.perf_counter()
        B.solve(u0s, ts, fetch=False, graphs=graphs)
        t2 = time.perf_counter()
        per_k = [B.snapshots(k) for k in range(len(ts))]  # every instance's snapshot k at once
        snaps = [[per_k[k][i] for k in range(len(ts))] for i in range(len(table))]
        t3 = time.perf_counter()
        return snaps, dict(B.stats), (t3 - t0, t1 - t0, t2 - t1, t3 - t2)

    batched()  # warm-up (code paths, caches)
    snaps, bstats, (b_total, b_setup, b_solve, b_fetch) = batched()
    say(f"c4: batched {b_total:.2f}s (setup {b_setup:.2f}, solve {b_solve:.2f}, download {b_fetch:.2f})")
    if not plain_leg:
        return {"batched_total_s": b_total, "batched_setup_s": b_setup, "batched_solve_s": b_solve,
                "batched_download_s": b_fetch, "batched_iterations": bstats["matvecs"] // 6,
                "batched_host_syncs": bstats["host_syncs"]}

    def plain(spec):
        A64, u0 = c4_instance(nh, spec, "fp64")
        o = L.nc_solve_opts()
        L.call("nc_solve_opts_default", C.byref(o))
        o.method, o.krylov_tol = L.NC_METHOD_KRYLOV, krylov_tol
        o.no_graphs = 0 if graphs else 1
        out = np.empty(len(ts) * len(u0))
        L.call("nc_solve_fixed", A64.handle, C.byref(o), L.ptr(u0, C.c_double), len(ts), L.ptr(ts, C.c_double),
               L.ptr(out, C.c_double), None)
        return out.reshape(len(ts), len(u0))

    plain(table[0])  # warm-up
    l1 = 0.0
    t0 = time.perf_counter()
    plain_out = []
    for i, spec in enumerate(table):
        plain_out.append(plain(spec))
        if i % 128 == 127:
            say(f"c4: plain fp64 {i + 1} instances, {time.perf_counter() - t0:.1f}s")
    p_total = time.perf_counter() - t0
    for i in range(len(table)):
        for k in range(len(ts)):
            l1 = max(l1, float(np.abs(snaps[i][k] - plain_out[i][k]).sum()))
    say(f"c4: plain fp64 {p_total:.2f}s")
    sizes = [t["states"] for t in table]
    return {
        "workload": f"512 instances (256 telegraph + 256 toggle, seeded table {C4_TABLE.name}), "
                    f"{min(sizes)}..{max(sizes)} states, {sum(sizes)} in total, t = 0:1:10; batched DP5(4) rtol 1e-8 "
                    f"in FP32 storage vs plain FP64 Krylov expv (m=30, tol {krylov_tol:g}) one instance at a time",
        "clock": "upload of every state space and generator to the download of every snapshot, both legs",
        "batched_setup": "nc_batch_create_models: spaces and generators built on 4 library worker threads; the "
                         "plain leg builds each instance's space and generator before its solve, one at a time",
        "batched_solves_per_s": len(table) / b_total,
        "batched_total_s": b_total,
        "batched_setup_s": b_setup,
        "batched_solve_s": b_solve,
        "batched_download_s": b_fetch,
        "batched_resident_solves_per_s": len(table) / b_solve,
        "batched_iterations": bstats["matvecs"] // 6,
        "batched_host_syncs": bstats["host_syncs"],
        "batched_steps_total": bstats["steps"],
        "batched_rejected_total": bstats["rejected"],
        "plain_fp64_solves_per_s": len(table) / p_total,
        "plain_fp64_total_s": p_total,
        "plain_method": "krylov",
        "plain_sample": "all 512 instances, one after another",
        "speedup_batched_vs_plain_fp64": p_total / b_total,
        "max_l1_batched_fp32_vs_plain_fp64": l1,
    }


C5_THETA = (10.0, 2.0, 1.0, 0.5, 0.2, 0.1)
MLP_FLOP_FORWARD = 2 * (8 * 64 + 64 * 64 + 64)  # 9,344 per sample
BF16_DENSE_PEAK_TFLOPS = 2500.0


def cascade_model(nh, theta):
    """0 -> A k0; A -> A + B k1 A; B -> B + C k2 B; A, B, C -> 0 at k3 A, k4 B, k5 C."""
    S = np.array([[1, 0, 0, -1, 0, 0], [0, 1, 0, 0, -1, 0], [0, 0, 1, 0, 0, -1]])
    props = [nh.propensity(lambda x, p: p[0] + 0.0 * x[0]), nh.propensity(lambda x, p: p[1] * x[0]),
             nh.propensity(lambda x, p: p[2] * x[1]), nh.propensity(lambda x, p: p[3] * x[0]),
             nh.propensity(lambda x, p: p[4] * x[1]), nh.propensity(lambda x, p: p[5] * x[2])]
    return nh.CmeModel(S, props, [float(v) for v in theta])


def _l1_between(a, b):
    """l1 distance of two sparse vectors over the union of their states (each state list
    holds every state once)."""
    sa, sb = np.asarray(a.states, dtype=np.int64), np.asarray(b.states, dtype=np.int64)
    ns = sa.shape[1]
    m = int(max(sa.max(initial=0), sb.max(initial=0))) + 1
    if sa.min(initial=0) < 0 or sb.min(initial=0) < 0 or m ** ns >= 2 ** 62:
        da = {tuple(x): v for x, v in zip(sa.tolist(), a.values)}
        db = {tuple(x): v for x, v in zip(sb.tolist(), b.values)}
        return float(sum(abs(da.get(x, 0.0) - db.get(x, 0.0)) for x in set(da) | set(db)))

    def key(s):
        k = np.zeros(s.shape[0], dtype=np.int64)
        for j in range(ns):
            k = k * m + s[:, j]
        return k

    _, inv = np.unique(np.concatenate([key(sa), key(sb)]), return_inverse=True)
    d = np.bincount(inv, weights=np.concatenate([a.values, -np.asarray(b.values)]))
    return float(np.abs(d).sum())


C5_PLAIN_LEVELS = (2, 6, 12, 24, 48)


def c5_rk(nh, prec):
    """C5's DP5(4) settings"""
    return nh.RK45(rtol=1e-6, atol=1e-10, precision=prec)


def c5_train(nh, T_rec=1.5, n_scalings=64, say=lambda m: None):
    """The learned adapter's network: records from plain SelectiveRStepAdapter(2, 2) solves
    of the cascade over n_scalings seeded parameter scalings, then device training."""
    rng = np.random.default_rng(0)
    rec = nh.TrainingRecords(capacity=1_000_000, per_adaptation=1024, label_mass=1e-10)
    t0 = time.perf_counter()
    for _ in range(n_scalings):
        th = np.asarray(C5_THETA) * np.exp(rng.uniform(np.log(0.5), np.log(2.0), size=6))
        alg = nh.AdaptiveFspSparse(ode_method=c5_rk(nh, "fp32"), space_adapter=nh.SelectiveRStepAdapter(2, 2, False))
        nh.solve(cascade_model(nh, th), nh.FspVectorSparse([[0, 0, 0]], [1.0]), (0.0, T_rec), alg,
                 saveat=np.linspace(0.0, T_rec, 6), fsptol=1e-4, records=rec)
    gen_s = time.perf_counter() - t0
    X, Y = rec.arrays()
    say(f"c5: {len(X)} records in {gen_s:.1f}s")
    t0 = time.perf_counter()
    net, loss = nh.train_adapter(X, Y, epochs=20, batch=4096, lr=2e-3, seed=0)
    train_s = time.perf_counter() - t0
    acc = float(np.mean((net.forward(X) > 0) == (Y > 0.5)))
    say(f"c5: trained in {train_s:.1f}s, loss {loss:.4f}, accuracy {acc:.4f}")
    return net, (X, Y, gen_s, train_s, loss, acc)


def c5_solve_only(nh, net_file, learned_levels=8, log=None):
    """Two learned FP32 solves of C5 with a network saved by an earlier call (which trains
    it and stores its parameters in net_file): the profiler view of one learned solve."""
    say = log or (lambda m: None)
    path = Path(net_file)
    if not path.exists():
        net, _ = c5_train(nh, say=say)
        np.save(path, net.params)
        return {"trained": str(path)}
    net = nh.AdapterMlp(0)
    net.params = np.load(path)
    model = cascade_model(nh, C5_THETA)
    res = []
    for _ in range(2):
        alg = nh.AdaptiveFspSparse(ode_method=c5_rk(nh, "fp32"),
                                   space_adapter=nh.LearnedAdapter(net, 2, learned_levels, False, threshold=0.3))
        t0 = time.perf_counter()
        out = nh.solve(model, nh.FspVectorSparse([[0, 0, 0]], [1.0]), (0.0, 5.0), alg,
                       saveat=np.linspace(0.0, 5.0, 6), fsptol=1e-4)
        for p in out.p:
            p.values
        res.append({"solve_s": time.perf_counter() - t0, **out.stats})
    return {"learned_fp32_solves": res}


def c5_learned(nh, T=5.0, T_rec=1.5, n_scalings=64, plain_levels=C5_PLAIN_LEVELS, learned_levels=8, log=None):
    """Config C5: the cascade from (0,0,0) with fsptol 1e-4 over [0, T], adaptive FSP with
    the learned adapter against plain SelectiveRStepAdapter(2, R), each at its best
    expansion depth (the learned adapter expands learned_levels levels from the states it
    keeps; the plain FP64 baseline is the fastest of plain_levels). The network is trained
    on the device on records from plain runs over n_scalings seeded parameter scalings. The
    plain depths are swept until the fastest lies inside the range (bracketed)."""
    def solve(model, adapter, ode, tend, records=None):
        alg = nh.AdaptiveFspSparse(ode_method=ode, space_adapter=adapter)
        t0 = time.perf_counter()
        out = nh.solve(model, nh.FspVectorSparse([[0, 0, 0]], [1.0]), (0.0, tend), alg,
                       saveat=np.linspace(0.0, tend, 6), fsptol=1e-4, records=records)
        return out, time.perf_counter() - t0

    rk = lambda prec: c5_rk(nh, prec)
    say = log or (lambda m: None)
    net, (X, Y, gen_s, train_s, loss, acc) = c5_train(nh, T_rec, n_scalings, say)

    model = cascade_model(nh, C5_THETA)
    learned = nh.LearnedAdapter(net, 2, learned_levels, False, threshold=0.3)
    solve(model, learned, rk("fp32"), T)  # warm-up
    out_l, t_l = solve(model, learned, rk("fp32"), T)
    out_l64, t_l64 = solve(model, learned, rk("fp64"), T)
    say(f"c5: learned {t_l:.2f}s (fp64 {t_l64:.2f}s)")
    plain_runs = {}
    for lv in plain_levels:
        out_p, t_p = solve(model, nh.SelectiveRStepAdapter(2, lv, False), rk("fp64"), T)
        plain_runs[f"rk45 R={lv}"] = (out_p, t_p)
        say(f"c5: plain fp64 rk45 R={lv} {t_p:.2f}s")
    l1_prec = max(_l1_between(a, b) for a, b in zip(out_l.p, out_l64.p))
    if plain_runs:
        best = min(plain_runs, key=lambda m: plain_runs[m][1])
        out_p, t_p = plain_runs[best]
        l1_vs_plain = max(_l1_between(a, b) for a, b in zip(out_l64.p, out_p.p))
        keys = list(plain_runs)
        bracketed = 0 < keys.index(best) < len(keys) - 1
    else:
        best, out_p, t_p, l1_vs_plain, bracketed = None, None, float("nan"), float("nan"), False
    batch = c5_batched(nh, net, learned_levels, best, T=T, log=say) if plain_runs else None
    fwd_ms = net.bench(n=1 << 21, reps=20, train=False)
    trn_ms = net.bench(n=1 << 21, reps=10, train=True)
    fwd_tf = MLP_FLOP_FORWARD * (1 << 21) / (fwd_ms * 1e-3) / 1e12
    trn_tf = 3 * MLP_FLOP_FORWARD * (1 << 21) / (trn_ms * 1e-3) / 1e12
    st_l = out_l.stats
    st_p = out_p.stats if out_p is not None else {"max_states": None, "adaptations": None}
    return {
        "workload": f"cascade (0->A 10, A->A+B 2A, B->B+C B, decay 0.5A 0.2B 0.1C) from (0,0,0), fsptol 1e-4, "
                    f"t in [0, {T:g}] with 6 outputs; LearnedAdapter(2, {learned_levels}, threshold 0.3) in FP32 vs "
                    f"plain SelectiveRStepAdapter(2, R, false) in FP64, R in {list(plain_levels)}, each with device "
                    f"DP5(4) rtol 1e-6",
        "training": {"records": int(len(X)), "positive_fraction": float(Y.mean()), "scalings": n_scalings,
                     "record_horizon": T_rec, "label_mass": 1e-10, "record_s": gen_s, "train_s": train_s,
                     "epochs": 20, "final_loss": loss, "train_accuracy": acc},
        "learned_fp32_rk45_s": t_l,
        "learned_fp64_rk45_s": t_l64,
        "learned_solves_per_s": 1.0 / t_l,
        "plain_fp64_solves_per_s": 1.0 / t_p,
        "plain_fp64_best": best,
        "plain_fp64_times_s": {m: v[1] for m, v in plain_runs.items()},
        "speedup_learned_vs_plain_fp64": t_p / t_l,
        "speedup_learned_fp64_vs_plain_fp64": t_p / t_l64,
        "plain_fp64_minimum_bracketed": bracketed,
        "speedup_learned_vs_plain_fp64_R2": (plain_runs["rk45 R=2"][1] / t_l) if "rk45 R=2" in plain_runs else None,
        "learned_max_states": st_l["max_states"],
        "plain_max_states": st_p["max_states"],
        "learned_adaptations": st_l["adaptations"],
        "learned_fallbacks": st_l["fallback_adaptations"],
        "plain_adaptations": st_p["adaptations"],
        "final_sink_mass_learned": float(out_l.sinks[-1].sum()),
        "final_sink_mass_plain": float(out_p.sinks[-1].sum()) if out_p is not None else None,
        "max_l1_learned_fp32_vs_fp64": l1_prec,
        "max_l1_learned_vs_plain_fp64": l1_vs_plain,
        "batched": batch,
        "mlp_roofline": {
            "bound": "mfma", "unit": "TFLOP/s", "peak": BF16_DENSE_PEAK_TFLOPS,
            "forward_achieved": fwd_tf, "forward_frac": fwd_tf / BF16_DENSE_PEAK_TFLOPS,
            "train_achieved": trn_tf, "train_frac": trn_tf / BF16_DENSE_PEAK_TFLOPS,
            "flop_per_sample_forward": MLP_FLOP_FORWARD, "samples_per_launch": 1 << 21,
            "forward_launch_ms": fwd_ms, "train_launch_ms": trn_ms,
        },
    }


C5_BATCH_INSTANCES = 16
C5_BATCH_WORKERS = 4


def c5_batched(nh, net, learned_levels, best, T=5.0, n_inst=C5_BATCH_INSTANCES, workers=C5_BATCH_WORKERS,
               log=None):
    """The batched + learned form on C5: a sweep of n_inst parameter points around C5's
    (every rate scaled by its own factor from [0.8, 1.25], seed 1; the network was trained
    on seed 0's scalings from [0.5, 2] over a shorter horizon), solved by
    nc_solve_adaptive_batch with the learned adapter, against the plain FP64 single-solve
    path (SelectiveRStepAdapter at the best depth of the bracketed sweep, one solve after
    another). Both legs run on one clock: model upload to the download of every snapshot's
    states and values. (Over t = 5 the full [0.5, 2] range reaches spaces of 10^8 states
    at its corners: a faster production and slower decay of C grow it without bound.)"""
    say = log or (lambda m: None)
    rng = np.random.default_rng(1)
    thetas = [np.asarray(C5_THETA) * np.exp(rng.uniform(np.log(0.8), np.log(1.25), size=6)) for _ in range(n_inst)]
    rk = lambda prec: nh.RK45(rtol=1e-6, atol=1e-10, precision=prec)
    p0 = nh.FspVectorSparse([[0, 0, 0]], [1.0])
    saveat = np.linspace(0.0, T, 6)
    plain_R = int(best.split("R=")[1])

    def fetch(outs):
        for o in outs:
            for p in o.p:
                p.states, p.values  # download
        return outs

    def batched(prec, nworkers):
        alg = nh.AdaptiveFspSparse(ode_method=rk(prec),
                                   space_adapter=nh.LearnedAdapter(net, 2, learned_levels, False, threshold=0.3))
        t0 = time.perf_counter()
        outs = fetch(nh.solve_adaptive_batch([cascade_model(nh, th) for th in thetas], p0, (0.0, T), alg,
                                             saveat=saveat, fsptol=1e-4, workers=nworkers))
        return outs, time.perf_counter() - t0

    def singles(adapter, prec, tag):
        alg = nh.AdaptiveFspSparse(ode_method=rk(prec), space_adapter=adapter)
        t0 = time.perf_counter()
        outs = []
        for i, th in enumerate(thetas):
            outs.append(fetch([nh.solve(cascade_model(nh, th), p0, (0.0, T), alg, saveat=saveat, fsptol=1e-4)])[0])
            if i % 4 == 3:
                say(f"c5 batch: {tag} {i + 1} instances, {time.perf_counter() - t0:.1f}s")
        return outs, time.perf_counter() - t0

    out_s, t_s = singles(nh.LearnedAdapter(net, 2, learned_levels, False, threshold=0.3), "fp32", "learned single")
    say("c5 batch: learned states per instance " + " ".join(str(o.stats["max_states"]) for o in out_s))
    batched("fp32", workers)  # warm-up of the worker threads' caches
    out_b, t_b = batched("fp32", workers)
    say(f"c5 batch: batched learned fp32 {t_b:.2f}s")
    out_b64, t_b64 = batched("fp64", workers)
    say(f"c5 batch: batched learned fp64 {t_b64:.2f}s")
    out_p, t_p = singles(nh.SelectiveRStepAdapter(2, plain_R, False), "fp64", f"plain fp64 R={plain_R}")
    same = all(np.array_equal(a.p[k].values, b.p[k].values) and np.array_equal(a.p[k].states, b.p[k].states)
               for a, b in zip(out_b, out_s) for k in range(len(a.p)))
    l1_prec = max(_l1_between(a.p[k], b.p[k]) for a, b in zip(out_b, out_b64) for k in range(len(a.p)))
    l1_plain = max(_l1_between(a.p[k], b.p[k]) for a, b in zip(out_b64, out_p) for k in range(len(a.p)))
    sink_ok = all(o.sinks[k].sum() <= 1e-4 * o.t[k] / T + 1e-13 for o in out_b + out_b64 + out_p
                  for k in range(len(o.t)))
    return {
        "workload": f"{n_inst} parameter points around C5 (each rate x exp(U[log 0.8, log 1.25]), seed 1), "
                    f"fsptol 1e-4, t in [0, {T:g}] with 6 outputs",
        "clock": "model upload to the download of every snapshot, both legs",
        "workers": workers,
        "batched_learned_fp32_s": t_b,
        "batched_learned_fp64_s": t_b64,
        "single_learned_fp32_s": t_s,
        "plain_fp64_single_s": t_p,
        "plain_adapter": f"SelectiveRStepAdapter(2, {plain_R}, false), the best depth of the C5 sweep",
        "batched_solves_per_s": n_inst / t_b,
        "batched_fp64_solves_per_s": n_inst / t_b64,
        "plain_fp64_solves_per_s": n_inst / t_p,
        "speedup_batched_learned_vs_plain_fp64": t_p / t_b,
        "speedup_batched_learned_fp64_vs_plain_fp64": t_p / t_b64,
        "speedup_batching_alone": t_s / t_b,
        "batched_equals_single_solves": bool(same),
        "sink_bound_held_every_snapshot": bool(sink_ok),
        "max_l1_batched_fp32_vs_fp64": l1_prec,
        "max_l1_batched_fp64_vs_plain_fp64": l1_plain,
        "max_states_per_instance": [int(o.stats["max_states"]) for o in out_b],
        "plain_max_states_per_instance": [int(o.stats["max_states"]) for o in out_p],
        "learned_adaptations": int(sum(o.stats["learned_adaptations"] for o in out_b)),
        "fallback_adaptations": int(sum(o.stats["fallback_adaptations"] for o in out_b)),
    }


def spmv_probe(nh, precision, reps):
    """Mat-vec launches plus a copy kernel of known size, for rocprofv3 --pmc runs."""
    import torch

    model = toggle_model(nh)
    space = nh.StateSpaceSparse(model.stoich_matrix, box_states(1024))
    A = nh.FspMatrixSparse(space, model.propensities, model.parameters, precision=precision, model=model)
    times = {}
    # every launch shape of the three modes (the counter rows carry the kernel's template
    # arguments; the last one is 3 for computed values, 4 for the matrix-free form)
    names = {0: "stored", 1: "computed_values", 2: "matrix_free"}
    default = A.info()["matrix_free"]
    for mode in (1, 2, 0):
        A.set_matrix_free(mode)
        for rows in ((1, 2) if mode else (1, 2, 4)):
            A.set_rows_per_thread(rows)
            times[f"{names[mode]} rows_per_thread={rows}"] = A.bench_matvec(reps=reps) * 1e3
    A.set_matrix_free(default)
    A.set_rows_per_thread(2)
    # the mat-vec with its Krylov projection epilogue over the basis sizes of one Krylov
    # step at m = 30 (1..30 vectors for the columns, 1 for the norm of A v_m), reps
    # launches each
    proj = {nv: A.bench_projection(nv, reps=reps) * 1e3 for nv in range(1, 31)}
    proj_step_mean = (sum(proj.values()) + proj[1]) / 31
    x = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")  # 256 MiB
    y = torch.empty_like(x)
    for _ in range(3):
        torch.mul(x, 1.0, out=y)  # reads 256 MiB, writes 256 MiB
    torch.cuda.synchronize()
    info = A.info()
    print(json.dumps({"probe": "spmv", "precision": precision, "reps": reps, "mean_launch_us": times,
                      "projection_mean_launch_us": proj, "projection_step_mix_mean_us": proj_step_mean,
                      "csr_bytes_per_matvec": info["csr_bytes_per_matvec"],
                      "own_bytes_per_matvec_default_mode": info["bytes_per_matvec"],
                      "copy_bytes_each_way": x.numel() * 4}), flush=True)


def sens_probe(nh, reps):
    """C3's sensitivity mat-vec launched reps times back to back in FP32 and in FP64 storage
    (t = 300, inside the pulse), plus the same copy kernel of known size as spmv_probe, for
    rocprofv3 --pmc runs."""
    import torch

    model = gene4_model(nh)
    space = nh.StateSpaceSparse(model.get_stoich_matrix(), gene4_states(255))
    res = {}
    for prec in ("fp32", "fp64"):
        SA = nh.ForwardSensFspMatrixSparse(model, space, precision=prec)
        res[prec] = {"mean_launch_us": SA.bench_matvec(reps=reps, t=300.0) * 1e3,
                     "own_bytes_per_matvec": SA.bytes_per_matvec}
        del SA
    x = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")  # 256 MiB
    y = torch.empty_like(x)
    for _ in range(3):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "sens", "reps": reps, **res, "copy_bytes_each_way": x.numel() * 4}), flush=True)


def c2_inputs():
    """C2's initial distribution as host arrays: the 1024^2 state list and delta(0, 0). They
    are the caller's data, made once before any clock starts."""
    states = box_states(1024)
    p0 = np.zeros(len(states))
    p0[0] = 1.0
    return states, p0


def c2_end_to_end(nh, precision, krylov_tol, ts, inputs):
    """One complete C2 solve from model upload to snapshot download: trace and upload the
    model, hand over the initial distribution, build the state space, generator and solver,
    integrate, download every snapshot."""
    model = toggle_model(nh)
    states, p0 = inputs
    out = nh.solve(model, nh.FspVectorSparse(states, p0, copy=False), (0.0, float(ts[-1])),
                   nh.KrylovExpv(m=30, tol=krylov_tol, precision=precision), saveat=ts)
    # the download of every snapshot: (probabilities, sink values) per output time
    snaps = [(p.values, s) for p, s in zip(out.p, out.sinks)]
    return snaps, out.stats


def snap_l1(a, b):
    """l1 distance of two downloaded snapshots (probabilities and sinks)"""
    return float(np.abs(a[0] - b[0]).sum() + np.abs(np.asarray(a[1]) - np.asarray(b[1])).sum())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--krylov-tol", type=float, default=1e-9,
                    help="expv local error tolerance per unit time of the timed solves: the loosest setting "
                         "whose FP64 solve stays within 1e-8 (l1) of the exact one (the FP64 accuracy reference "
                         "always uses 1e-12)")
    ap.add_argument("--no-graphs", action="store_true", help="launch kernels eagerly instead of replaying graphs")
    ap.add_argument("--no-c3", action="store_true", help="skip the forward-sensitivity (C3) measurement")
    ap.add_argument("--no-c4", action="store_true", help="skip the batched (C4) measurement")
    ap.add_argument("--no-c5", action="store_true", help="skip the learned-adapter (C5) measurement")
    ap.add_argument("--c4-plain", default="all", choices=["all", "none"],
                    help="with --only c4: time the plain FP64 comparison too (none: the batched solves alone, "
                         "for profiler runs)")
    ap.add_argument("--c5-plain", default="all", choices=["all", "none"],
                    help="time the plain FP64 adapters for C5 (none: learned only, for profiler runs)")
    ap.add_argument("--spmv-probe", type=int, default=0, metavar="REPS",
                    help="only build the matrix, launch the mat-vec REPS times back to back and run a plain "
                         "256 MiB copy kernel three times (the calibration pair for HBM counter runs)")
    ap.add_argument("--sens-probe", type=int, default=0, metavar="REPS",
                    help="only build C3's sensitivity operator in FP32 and FP64, launch its mat-vec REPS times back "
                         "to back each, and run the 256 MiB copy kernel three times (counter runs)")
    ap.add_argument("--only", choices=["c3", "c4", "c5", "c5-solve", "mlp"],
                    help="run one leg alone and print its figures (for profiler runs); no contract line. "
                         "c5-solve: train C5's network and save it to --net-file, or, when that file exists, "
                         "load it and run two learned FP32 solves")
    ap.add_argument("--net-file", default="c5_net.npy", help="network parameters for --only c5-solve")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; runs on more than one GPU are not supported (how "
                                   "a solve or a batch would shard is documented with the library)",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    say = lambda m: print(m, file=sys.stderr, flush=True)

    import numcme_hip as nh
    from numcme_hip import _lib as L

    if args.spmv_probe:
        spmv_probe(nh, args.precision, args.spmv_probe)
        return
    if args.sens_probe:
        sens_probe(nh, args.sens_probe)
        return
    if args.only == "mlp":
        # the network kernels alone on 2,097,152 synthetic rows (HIP events per launch)
        net = nh.AdapterMlp(seed=0)
        fwd_ms = net.bench(n=1 << 21, reps=20, train=False)
        trn_ms = net.bench(n=1 << 21, reps=10, train=True)
        print(json.dumps({"mlp_forward_ms": fwd_ms, "mlp_train_ms": trn_ms,
                          "forward_TFLOPs": MLP_FLOP_FORWARD * (1 << 21) / (fwd_ms * 1e-3) / 1e12,
                          "train_TFLOPs": 3 * MLP_FLOP_FORWARD * (1 << 21) / (trn_ms * 1e-3) / 1e12}), flush=True)
        return
    if args.only == "c3":
        print(json.dumps(c3_sensitivity(nh, L)), flush=True)
        return
    if args.only == "c4":
        print(json.dumps(c4_batch(nh, L, args.krylov_tol, graphs=not args.no_graphs, log=say,
                                  plain_leg=args.c4_plain == "all")), flush=True)
        return
    if args.only == "c5-solve":
        print(json.dumps(c5_solve_only(nh, args.net_file, log=say)), flush=True)
        return
    if args.only == "c5":
        print(json.dumps(c5_learned(nh, plain_levels=C5_PLAIN_LEVELS if args.c5_plain == "all" else (), log=say)),
              flush=True)
        return

    ts = np.arange(0.0, 11.0)
    # FP64 reference at tolerance 1e-12 (accuracy check of the timed solves)
    inputs = c2_inputs()
    ref, _ = c2_end_to_end(nh, "fp64", 1e-12, ts, inputs)

    # timed region: K complete solves, each from model upload to snapshot download
    for _ in range(max(args.warmup, 0)):
        c2_end_to_end(nh, args.precision, args.krylov_tol, ts, inputs)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        snaps, stats = c2_end_to_end(nh, args.precision, args.krylov_tol, ts, inputs)
    elapsed = time.perf_counter() - t0
    l1 = max(snap_l1(snaps[k], ref[k]) for k in range(len(ts)))
    say(f"c2: {args.steps / elapsed:.3f} solves/s end to end, l1 vs fp64 {l1:.2e}")
    # the plain FP64 path on the same clock and tolerance
    t0 = time.perf_counter()
    snaps64, stats64 = c2_end_to_end(nh, "fp64", args.krylov_tol, ts, inputs)
    fp64_s = time.perf_counter() - t0
    l1_64 = max(snap_l1(snaps64[k], ref[k]) for k in range(len(ts)))

    # the same solves with inputs resident on the device (restart from the uploaded state)
    model = toggle_model(nh)
    space = nh.StateSpaceSparse(model.stoich_matrix, box_states(1024))
    t_build = time.perf_counter()
    A = nh.FspMatrixSparse(space, model.propensities, model.parameters, precision=args.precision, model=model)
    build_s = time.perf_counter() - t_build
    info = A.info()
    N = info["n_rows"]
    u0 = np.zeros(N)
    u0[0] = 1.0
    sol = Solver(L, A, args.precision, False, graphs=not args.no_graphs, tol=args.krylov_tol)
    sol.set_state(u0)
    sol.solve(ts)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sol.solve(ts)
    resident_s = (time.perf_counter() - t0) / args.steps
    sol.free()

    # mat-vec roofline: one more solve with HIP events recorded around every mat-vec
    # launch on the solver's stream (graphs off, since events cannot sit inside them):
    # mode 1 times the launches as the solves run them (the mat-vec with its projection
    # epilogue), mode 2 the plain mat-vec (the projection sums then run as a separate pass)
    def timed_solve(mode=1):
        tim = Solver(L, A, args.precision, mode, tol=args.krylov_tol)
        tim.set_state(u0)
        tim.solve(ts)
        st = tim.stats()
        tim.free()
        return (st["matvec_ms"] / st["matvecs_timed"]) * 1e-3 if st["matvecs_timed"] else float("nan"), \
            st["matvecs_timed"]

    mf_on = bool(info["matrix_free"])
    # the stored generator at 2 rows per thread (the default shape) carries the epilogue
    projection = not mf_on and info["rows_per_thread"] == 2
    avg_mv_s, mv_n = timed_solve(1)
    avg_plain_s, plain_n = timed_solve(2) if projection else (avg_mv_s, mv_n)
    back_to_back_ms = A.bench_matvec(reps=200)
    # the epilogue kernel back to back over one Krylov step's basis sizes (1..30, then 1)
    proj_b2b_ms = ((sum(A.bench_projection(nv, reps=20) for nv in range(1, 31)) + A.bench_projection(1, reps=20)) / 31
                   if projection else None)
    # the projection epilogue reads the column's basis vectors V_0..V_j at the state rows:
    # j + 1 = 1..m for the m columns and 1 for the norm of A v_m, m + 1 launches a step
    m_k = 30
    proj_bytes = (m_k * (m_k + 1) // 2 + 1) / (m_k + 1) * info["n_states"] * (4 if args.precision == "fp32" else 8)
    stored = None
    if mf_on:
        # the same kernel reading the stored generator, for comparison
        A.set_matrix_free(0)
        s_avg, _ = timed_solve(2)
        s_info = A.info()
        stored = {"avg_launch_us": s_avg * 1e6, "achieved_csr_equivalent": info["csr_bytes_per_matvec"] / s_avg / 1e9,
                  "bytes_per_launch_own_layout": s_info["bytes_per_matvec"],
                  "achieved_own_bytes": s_info["bytes_per_matvec"] / s_avg / 1e9,
                  "traffic": measured_traffic(info["rows_per_thread"], args.precision, 0)}
        A.set_matrix_free(info["matrix_free"])
    rows_tuned, _ = A.autotune(reps=30)  # reported only: the timed solves use the default shape
    A.set_rows_per_thread(info["rows_per_thread"])
    # the dominant kernel of the timed solves: the mat-vec with its projection epilogue
    # (CSR-equivalent mat-vec bytes plus the basis vectors it reads), or the plain mat-vec
    launch_bytes = info["csr_bytes_per_matvec"] + (proj_bytes if projection else 0.0)
    achieved = launch_bytes / avg_mv_s / 1e9
    spmv_achieved = info["csr_bytes_per_matvec"] / avg_plain_s / 1e9
    line = {
        "metric": "FSP solves/s at \u21131\u22641e-6 (batched+learned vs plain FP64); SpMV HBM GB/s",
        "value": args.steps / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-storage/f64-accumulate" if args.precision == "fp32" else "f64",
        "data": "synthetic",
        "config": {
            "workload": "toggle switch, fixed FSP space 0..1023^2 (1,048,576 states + 4 sinks, 5,240,832 nnz), "
                        "t=0..10 at 11 outputs, device Krylov expv m=30",
            "precision": args.precision,
            "clock": "model upload (trace, initial distribution, state space, generator, solver) to the download "
                     "of all 11 snapshots; the initial distribution's host arrays are made once before the clock",
            "value_is": "plain single-solve FP32-storage C2 solves/s at l1 <= 1e-6 (l1_vs_fp64_gpu); the batched "
                        "and learned forms are metric_components",
            "krylov_tol": args.krylov_tol,
            "krylov_steps_per_solve": stats["steps"],
            "matvecs_per_solve": stats["matvecs"],
        },
        "l1_vs_fp64_gpu": l1,
        "fp64_solves_per_s": 1.0 / fp64_s,
        "fp64_l1_vs_tol_1e-12": l1_64,
        "fp64_krylov_steps_per_solve": stats64["steps"],
        "resident_solves_per_s": 1.0 / resident_s,
        "matrix_build_s": build_s,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": measured_traffic(info["rows_per_thread"], args.precision, info["matrix_free"], projection),
            "traffic_source": (PROJECTION_PROFILE if projection else TRAFFIC_PROFILE).name,
            "kernel": ("k_matvec with the Arnoldi projection epilogue (FSP SpMV, stored generator, plus the "
                       "column's projection sums over 512-row blocks)" if projection else
                       "k_matvec (FSP SpMV), " + ("computed values: off-diagonal values from per-reaction tables, "
                                                  "columns from row offsets, the FP64 diagonal read; achieved "
                                                  "counts the CSR-equivalent bytes" if mf_on else "stored generator")),
            "bytes_per_launch": launch_bytes,
            "bytes_per_launch_csr": info["csr_bytes_per_matvec"],
            "bytes_per_launch_projection": proj_bytes if projection else 0.0,
            "bytes_per_launch_own_layout": info["bytes_per_matvec"],
            "stored_generator": stored,
            "avg_launch_us": avg_mv_s * 1e6,
            "launches_timed": mv_n,
            "back_to_back_step_mix_launch_us": proj_b2b_ms * 1e3 if projection else None,
            "spmv": {
                "kernel": "k_matvec (FSP SpMV), stored generator, timed in a solve whose projection sums run as "
                          "a separate pass" if projection else "as above",
                "avg_launch_us": avg_plain_s * 1e6,
                "launches_timed": plain_n,
                "achieved_csr_equivalent": spmv_achieved,
                "frac": spmv_achieved / HBM_PEAK_GBPS,
                "achieved_own_bytes": info["bytes_per_matvec"] / avg_plain_s / 1e9,
                "traffic": measured_traffic(info["rows_per_thread"], args.precision, info["matrix_free"]),
                "traffic_source": TRAFFIC_PROFILE.name,
                "back_to_back_launch_us": back_to_back_ms * 1e3,
            },
            "rows_per_thread": info["rows_per_thread"],
            "autotuned_rows_per_thread": rows_tuned,
            "computed_col_slots": info["computed_col_slots"],
        },
        "cpu_baseline": None,
    }
    comps = {}
    if not args.no_c3:
        line["c3_sensitivity"] = c3_sensitivity(nh, L)
    if not args.no_c4:
        c4 = c4_batch(nh, L, args.krylov_tol, graphs=not args.no_graphs, log=say)
        line["c4_batch"] = c4
        comps["batched_vs_plain_fp64"] = c4["speedup_batched_vs_plain_fp64"]
        comps["batched_solves_per_s"] = c4["batched_solves_per_s"]
    if not args.no_c5:
        plain = C5_PLAIN_LEVELS if args.c5_plain == "all" else ()
        c5 = c5_learned(nh, plain_levels=plain, log=say)
        line["c5_learned"] = c5
        comps["learned_vs_plain_fp64"] = c5["speedup_learned_vs_plain_fp64"]
        comps["learned_fp64_vs_plain_fp64"] = c5["speedup_learned_fp64_vs_plain_fp64"]
        if c5["batched"]:
            comps["batched_learned_vs_plain_fp64"] = c5["batched"]["speedup_batched_learned_vs_plain_fp64"]
            comps["batched_learned_fp64_vs_plain_fp64"] = c5["batched"]["speedup_batched_learned_fp64_vs_plain_fp64"]
            comps["batched_learned_solves_per_s"] = c5["batched"]["batched_solves_per_s"]
    comps["spmv_hbm_GBps"] = spmv_achieved
    line["metric_components"] = comps
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(model.parameters, args.krylov_tol, stats64["steps"], log=say)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
