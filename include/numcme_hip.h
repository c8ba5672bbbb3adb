/*
 * numcme_hip.h — C ABI of the MI355X-native finite-state-projection (FSP) engine.
 *
 * The library replaces, for one GPU, the FSP hot path of NumCME.jl: the sparse state
 * space, the FSP generator and its mat-vec, the time integrator with the sink-mass
 * budget, and the R-step space adapters. Each entry point below names the reference
 * code it replaces (paths relative to the NumCME.jl repository).
 *
 * Conventions (all entry points):
 *  - Every function returns an int status (NC_OK == 0). On failure a message is
 *    available from nc_last_error() in the calling thread. No C++ exception crosses
 *    the ABI.
 *  - Handles are opaque and owned by the library; destroy them with *_destroy.
 *  - Host pointers are borrowed for the duration of the call only. Outputs are copied
 *    into caller-allocated arrays whose sizes are stated per function.
 *  - Indices are 0-based. "None" is -1 (state connectivity) and reaction numbers are
 *    0-based (sink connectivity). A 1-based host shim adds 1 to every index.
 *  - Each handle owns one HIP stream (default priority). Handles are not thread-safe;
 *    calls on one handle serialise, as in the single-threaded reference.
 *  - The library needs a HIP device; with none it fails with NC_ERR_NO_DEVICE. There
 *    is no CPU fallback.
 */
#ifndef NUMCME_HIP_H
#define NUMCME_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------------------- */
enum {
  NC_OK = 0,
  NC_ERR_ARG = 1,         /* invalid argument (the reference throws ArgumentError)        */
  NC_ERR_HIP = 2,         /* HIP runtime error                                            */
  NC_ERR_ALLOC = 3,       /* device or host allocation failed                             */
  NC_ERR_OVERFLOW = 4,    /* a size or index exceeds the library's 32-bit limits          */
  NC_ERR_UNSUPPORTED = 5, /* feature outside what the library supports                  */
  NC_ERR_NUMERIC = 6,     /* integrator failure (step size underflow, too many steps)     */
  NC_ERR_INTERNAL = 7,
  NC_ERR_NO_DEVICE = 8
};

/* Storage precision of matrix values and probability vectors. */
enum { NC_FP64 = 0, NC_FP32 = 1 };

/* Copy the calling thread's last error message into buf (always NUL-terminated). */
int nc_last_error(char* buf, size_t len);
/* Library version and the HIP device it runs on. */
int nc_version(int32_t* major, int32_t* minor);
/* Return the device blocks the library keeps cached for reuse (freed objects' memory) to
 * the driver; synchronises the device. No reference counterpart (Julia's GC plays this
 * part for the reference's arrays). */
int nc_memory_trim(void);
int nc_device_name(char* buf, size_t len);

/* ======================================================================================
 * State space — replaces StateSpaceSparse (src/statespace/sparse/sparsestatespace.jl).
 * States are int32 vectors of ns species; a state space holds at most 2^31-2 states
 * and at most 32 reactions.
 * ==================================================================================== */
typedef struct nc_space_s* nc_space;

/* StateSpaceSparse(stoich_matrix, initstates) — sparsestatespace.jl:103-124 (and the
 * single-state form :142-144). stoich: ns x nr, column-major (column r = s_r).
 * init_states: n_init x ns, row-major. Duplicates and states with a negative entry are
 * dropped, first occurrence wins, exactly as _addstates! (:208-267). */
int nc_space_create(int32_t ns, int32_t nr, const int32_t* stoich, int64_t n_init,
                    const int32_t* init_states, nc_space* out);
int nc_space_destroy(nc_space space);
/* expand!(space, levels; onlyreactions) — sparsestatespace.jl:153-194. only_reactions
 * holds n_only 0-based reaction numbers; n_only == 0 means all reactions. The new states
 * get the indices the reference assigns (LIFO exploration, first occurrence wins). */
int nc_space_expand(nc_space space, int32_t levels, int32_t n_only, const int32_t* only_reactions);
/* deleteat!(space, ids) — sparsestatespace.jl:276-331. ids: n_ids 0-based indices,
 * strictly increasing (the reference requires sorted unique ids). */
int nc_space_delete(nc_space space, int64_t n_ids, const int64_t* ids);
/* get_state_count / get_sink_count — sparsestatespace.jl:55,62 */
int nc_space_count(nc_space space, int64_t* n_states, int32_t* n_sinks);
/* get_states — sparsestatespace.jl:69. out: n_states x ns int32, row-major. */
int nc_space_get_states(nc_space space, int32_t* out);
/* get_state_connectivity — sparsestatespace.jl:78. out: n_states x nr int32; entry (i,r)
 * is the index of x_i - s_r, or -1. */
int nc_space_get_state_conn(nc_space space, int32_t* out);
/* get_sink_connectivity — sparsestatespace.jl:80. out: n_states x nr int32; entry (i,r)
 * is r when x_i + s_r is non-negative and absent from the space, else -1. */
int nc_space_get_sink_conn(nc_space space, int32_t* out);
/* get_statedict lookups — sparsestatespace.jl:76. states: n x ns row-major; out: index
 * or -1 per state. */
int nc_space_lookup(nc_space space, int64_t n, const int32_t* states, int64_t* out);

/* ======================================================================================
 * Model — replaces CmeModel and the propensity kinds (src/cmemodel/cmemodel.jl:49-53,
 * src/cmemodel/propensity.jl:50-153). Propensities are postfix programs evaluated in
 * FP64 on a stack of depth NC_PROG_STACK.
 * ==================================================================================== */
#define NC_PROG_STACK 16
enum nc_opcode {
  NC_OP_CONST = 0,  /* push c                              */
  NC_OP_PARAM = 1,  /* push theta[arg]                     */
  NC_OP_SPECIES = 2,/* push x[arg]                         */
  NC_OP_TIME = 3,   /* push t                              */
  NC_OP_ADD = 4, NC_OP_SUB = 5, NC_OP_MUL = 6, NC_OP_DIV = 7, NC_OP_POW = 8,
  NC_OP_NEG = 9, NC_OP_EXP = 10, NC_OP_LOG = 11, NC_OP_SIN = 12, NC_OP_COS = 13,
  NC_OP_SQRT = 14, NC_OP_ABS = 15,
  NC_OP_MAX = 16,   /* max(a,b) = (b < a) ? a : b, as Julia's Base.max  */
  NC_OP_MIN = 17,   /* min(a,b) = (b < a) ? b : a                      */
  NC_OP_LT = 18, NC_OP_LE = 19, NC_OP_GT = 20, NC_OP_GE = 21, /* push 1.0 or 0.0     */
  NC_OP_SELECT = 22 /* pops c, a, b (b on top): push (c != 0) ? a : b */
};
typedef struct {
  int32_t op;
  int32_t arg;
  double c;
} nc_instr;

/* Propensity kinds — propensity.jl:50-52 (StandardTimeInvariantPropensity), :91-94
 * (SeparableTimeVaryingPropensity), :122-124 (JointTimeVaryingPropensity). */
enum { NC_PROP_TIME_INVARIANT = 0, NC_PROP_SEPARABLE = 1, NC_PROP_JOINT = 2 };

typedef struct nc_model_s* nc_model;
/* CmeModel(stoich_matrix, propensities, parameters) — cmemodel.jl:49-53.
 * kinds[r] is the propensity kind of reaction r. prog_offsets has 2*nr+1 entries;
 * reaction r's "x program" is code[prog_offsets[2r] .. prog_offsets[2r+1]) and its
 * "t program" is code[prog_offsets[2r+1] .. prog_offsets[2r+2]):
 *   time-invariant: x program f(x, theta); t program empty
 *   separable:      x program statefactor(x, theta); t program tfactor(t, theta)
 *   joint:          x program f(t, x, theta); t program empty                        */
int nc_model_create(int32_t ns, int32_t nr, const int32_t* stoich, int32_t n_params,
                    const double* params, const int32_t* kinds, const int64_t* prog_offsets,
                    const nc_instr* code, nc_model* out);
int nc_model_destroy(nc_model model);
/* Partial derivatives of the propensities — CmeModelWithSensitivity (cmemodel.jl:104-108,
 * 157-170) with PropensityGradient semantics (propensitygrad.jl:56-67). pattern is
 * nr x n_params row-major (1 where a_r depends on theta_p, the reaction x parameter
 * sparsity pattern). For pair k = r*n_params + p, dprog_offsets[2k..2k+2) delimit two
 * programs: d(x program)/d theta_p and d(t program)/d theta_p (the latter only for
 * separable reactions; empty otherwise). dprog_offsets has 2*nr*n_params+1 entries. */
int nc_model_set_gradients(nc_model model, const uint8_t* pattern, const int64_t* dprog_offsets,
                           const nc_instr* dcode);
/* Evaluate reaction r's x program (or, when which == 1, its t program) at n states
 * (row-major n x ns) and time t — the propensity call methods of propensity.jl:145-153.
 * Runs on the device; used to check the program evaluator. */
int nc_model_eval(nc_model model, int32_t r, int32_t which, double t, int64_t n,
                  const int32_t* states, double* out);

/* ======================================================================================
 * FSP matrix — replaces FspMatrixSparse (src/fspmatrix/sparse/fspsparsematrix.jl).
 * Rows/columns: the states of the space in order, then one sink per reaction.
 * ==================================================================================== */
typedef struct nc_matrix_s* nc_matrix;
typedef struct {
  int64_t n_rows;        /* states + sinks                                         */
  int64_t n_states;
  int32_t n_sinks;
  int32_t ell_width;     /* merged time-invariant off-diagonal slots per state row */
  int32_t n_tv;          /* time-varying (separable + joint) reactions             */
  int32_t precision;
  int64_t nnz_reference; /* nonzeros of the reference's sparse matrices (explicit
                            zeros included, duplicates merged per matrix)           */
  int64_t n_sink_entries;
  double bytes_per_matvec;   /* bytes the mat-vec kernel moves (its own layout)     */
  double csr_bytes_per_matvec; /* nnz*(s_v+4) + (N+1)*4 + 2*N*s_v, CSR-equivalent   */
  double norm1;          /* max column abs sum (2 * max |diagonal|)                 */
  int32_t rows_per_thread; /* mat-vec kernel: consecutive state rows per thread     */
  int32_t computed_col_slots; /* slots whose columns are row + constant offset in
                            every nonzero row, computed instead of loaded          */
  int32_t matrix_free;   /* mat-vec mode (nc_matrix_set_matrix_free): 0 stored arrays,
                            1 computed values with the stored diagonal, 2 matrix-free */
  int32_t matrix_free_available; /* bit 0: mode 1 available, bit 1: mode 2 available */
} nc_matrix_info_t;

/* FspMatrixSparse{RealT}(space, propensities; parameters) — fspsparsematrix.jl:47-108.
 * precision NC_FP64 stores FP64 values; NC_FP32 stores FP32 off-diagonal and sink
 * values with an FP64 diagonal equal to minus the exact sum of the FP32 column entries,
 * so every column of the stored generator sums to zero exactly. */
int nc_matrix_create(nc_space space, nc_model model, int32_t precision, nc_matrix* out);
int nc_matrix_destroy(nc_matrix A);
int nc_matrix_info(nc_matrix A, nc_matrix_info_t* info);
/* matvec!(out, t, A, v) (accumulate == 0) and matvecadd!(out, t, A, v) (accumulate != 0)
 * — fspsparsematrix.jl:196-247. v, out: host arrays of n_rows doubles. */
int nc_matrix_matvec(nc_matrix A, double t, const double* v, double* out, int32_t accumulate);
/* nc_matrix_matvec without any allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (tight serving loops, code that forbids
 * hipMalloc). The input, its FP32 copy and the output are staged in a fixed device array
 * of 48 MiB per device, which limits the generator to about 2.5M rows (FP32 storage) or
 * 3.1M rows (FP64); larger ones fail with NC_ERR_ARG. Every call on a device shares that
 * array, so calls that overlap on one device race; results equal nc_matrix_matvec's. */
int nc_matrix_matvec_noalloc(nc_matrix A, double t, const double* v, double* out, int32_t accumulate);
/* The generator A(t) as COO triplets (row, col, value) in the stored precision, for
 * inspection. Call with rows == NULL to get *nnz only. */
int nc_matrix_export(nc_matrix A, double t, int64_t* nnz, int64_t* rows, int64_t* cols, double* vals);
/* Repeat the mat-vec kernel `reps` times on device-resident vectors (the integrators' form:
 * stored-precision input, FP64 output) and report the mean kernel duration from HIP events
 * on the matrix's stream. No reference counterpart. */
int nc_matrix_bench_matvec(nc_matrix A, double t, int32_t reps, double* mean_ms);
/* The same for the mat-vec that the Krylov integrator runs with its projection epilogue:
 * out = A v plus the per-block projection sums of nv basis vectors (1..32), v the last of
 * them. NC_ERR_UNSUPPORTED for generators whose mat-vec has no such epilogue. No reference
 * counterpart. */
int nc_matrix_bench_projection(nc_matrix A, int32_t nv, int32_t reps, double* mean_ms);
/* Computed mat-vecs. A generator on a box state space in lattice order (species 0 fastest,
 * at most 4 species, at most 8 distinct stoichiometries with at most 4 reactions each)
 * whose propensities are all time-invariant and each read at most one species keeps,
 * besides its stored arrays, one table per reaction: the propensity over its species'
 * range, evaluated by the build's own evaluator. In the integrators' mat-vec (FP64 output;
 * other launches read the stored arrays) mode 1 then takes every off-diagonal value from
 * the tables and every column from a row offset, reading the stored FP64 diagonal (every
 * stoichiometry must hold one reaction changing one species). Mode 2 computes the
 * diagonal as well and reads only v and the sink lists (matrix-free). Both give results
 * bit-identical to the stored generator's (mode 0, the default: on MI355X it is the
 * fastest, the computed forms being bound by their index arithmetic rather than by
 * memory). Modes the generator does not qualify for fail with NC_ERR_UNSUPPORTED. No
 * reference counterpart (the reference always stores its generator,
 * fspsparsematrix.jl:47-108). */
int nc_matrix_set_matrix_free(nc_matrix A, int32_t mode);
/* Mat-vec launch shape: consecutive state rows per thread (1, 2 or 4). No reference
 * counterpart (the reference's CSC mul! has no launch shape). */
int nc_matrix_set_rows_per_thread(nc_matrix A, int32_t rows);
/* Time every launch shape on this matrix (`reps` launches each, every one after a read of
 * 512 MB that evicts the matrix from the caches, as a solve's vector passes do) and keep
 * the fastest; *best_ms receives its mean kernel duration. No reference counterpart. */
int nc_matrix_autotune(nc_matrix A, int32_t reps, double* best_ms);

/* ======================================================================================
 * Time integration — replaces the integrator loop of solve() (fspsolve.jl:10-41) with
 * a device-resident Krylov expv (time-invariant generators) or an adaptive
 * Dormand-Prince 5(4) Runge-Kutta (any generator).
 * ==================================================================================== */
enum { NC_METHOD_AUTO = 0, NC_METHOD_KRYLOV = 1, NC_METHOD_RK45 = 2 };
typedef struct {
  int32_t method;      /* NC_METHOD_*                                              */
  int32_t krylov_m;    /* Krylov subspace dimension (default 30)                   */
  double krylov_tol;   /* expv local error tolerance per unit time (default 1e-12) */
  double rtol, atol;   /* RK tolerances (defaults 1e-8, 1e-14)                     */
  double h0;           /* first RK step, 0 = automatic                             */
  int64_t max_steps;   /* 0 = unlimited                                            */
  int32_t time_kernels;/* record HIP events around every mat-vec launch: 1 as the
                          solve runs (Krylov: the mat-vec with its projection sums where
                          the generator allows), 2 plain mat-vecs (the projection sums by
                          a separate pass)                                          */
  int32_t no_graphs;  /* 1: launch kernels directly instead of replaying HIP graphs */
} nc_solve_opts;

typedef struct {
  int64_t steps, rejected, matvecs, host_syncs;
  int64_t matvecs_timed;
  double matvec_ms;    /* summed event-timed mat-vec kernel durations              */
  double error_estimate;
  int64_t adaptations; /* adaptive solves: state-space adaptations                 */
  int64_t learned_adaptations;  /* of which the learned adapter selected the states */
  int64_t fallback_adaptations; /* of which fell back to plain SelectiveRStep       */
  int64_t max_states;  /* adaptive solves: largest state count reached             */
} nc_solve_stats;

typedef struct nc_solver_s* nc_solver;
int nc_solve_opts_default(nc_solve_opts* opts);
int nc_solver_create(nc_matrix A, const nc_solve_opts* opts, nc_solver* out);
int nc_solver_destroy(nc_solver s);
/* Upload the ODE state u = [p; sinks] (n_rows doubles) at time t0. */
int nc_solver_set_state(nc_solver s, double t0, const double* u);
/* Reset the state to the one uploaded by the last nc_solver_set_state (kept on the
 * device), so repeated solves need no host transfer. */
int nc_solver_restart(nc_solver s);
/* Integrate from the current time through the ascending output times t_out[0..n_out),
 * keeping one snapshot per output time on the device (DE.solve with saveat,
 * fspsolve.jl:27). Returns when the device work is complete. */
int nc_solver_run(nc_solver s, int32_t n_out, const double* t_out);
/* Copy snapshot k (n_rows doubles) to the host. */
int nc_solver_get_snapshot(nc_solver s, int32_t k, double* u);
int nc_solver_get_state(nc_solver s, double* t, double* u);
int nc_solver_stats(nc_solver s, nc_solve_stats* st);

/* Fixed-space solve — fspsolve.jl:10-41. u0: n_rows doubles; u_out: n_out x n_rows. */
int nc_solve_fixed(nc_matrix A, const nc_solve_opts* opts, const double* u0, int32_t n_out,
                   const double* t_out, double* u_out, nc_solve_stats* st);

/* ======================================================================================
 * Forward sensitivities — replaces ForwardSensFspMatrixSparse
 * (src/forwardsensfspmatrix/forwardsensfspmatrixsparse/sensfspmatrixsparse.jl:31-142) and
 * the fixed-space form of the sensitivity solve (src/forwardsenscme/sparse/
 * forwardsenscmesparse.jl:99-215). Vectors hold n_params+1 blocks of n_rows entries,
 * block-major: [p; sinks; S_1; dsinks_1; ...] as the reference's u (:132-138).
 * ==================================================================================== */
typedef struct nc_sensmatrix_s* nc_sensmatrix;
/* ForwardSensFspMatrixSparse{RealT}(model, statespace) — sensfspmatrixsparse.jl:31-95. The
 * model needs its gradients (nc_model_set_gradients); at most 31 parameters and 128
 * (reaction, parameter) pairs. */
int nc_sensmatrix_create(nc_space space, nc_model model, int32_t precision, nc_sensmatrix* out);
int nc_sensmatrix_destroy(nc_sensmatrix SA);
/* Rows per block, parameter count, derivative pairs and the bytes one mat-vec moves. */
int nc_sensmatrix_info(nc_sensmatrix SA, int64_t* n_rows, int32_t* n_params, int32_t* n_pairs,
                       double* bytes_per_matvec);
/* matvec!(out, t, SA, vs) — sensfspmatrixsparse.jl:97-142. v, out: (n_params+1) x n_rows. */
int nc_sensmatrix_matvec(nc_sensmatrix SA, double t, const double* v, double* out);
/* Mean duration (ms) of `reps` back-to-back launches. No reference counterpart. */
int nc_sensmatrix_bench_matvec(nc_sensmatrix SA, double t, int32_t reps, double* mean_ms);
/* A solver over the sensitivity system; nc_solver_set_state / run / get_snapshot then take
 * (n_params+1) x n_rows vectors. */
int nc_solver_create_sens(nc_sensmatrix SA, const nc_solve_opts* opts, nc_solver* out);
/* Fixed-space sensitivity solve. u0: (n_params+1) x n_rows; u_out: n_out of those. */
int nc_solve_sens_fixed(nc_sensmatrix SA, const nc_solve_opts* opts, const double* u0, int32_t n_out,
                        const double* t_out, double* u_out, nc_solve_stats* st);

/* ======================================================================================
 * Batched solves — no reference counterpart (the reference solves one model per call).
 * Many time-invariant generators of uneven size are integrated together with
 * Dormand-Prince 5(4), per-instance step control, one replayed HIP graph per iteration.
 * ==================================================================================== */
typedef struct nc_batch_s* nc_batch;
/* Build a batch from n generators (all of one precision, time-invariant). The matrices are
 * copied; they may be destroyed afterwards. */
int nc_batch_create(int32_t n, const nc_matrix* mats, nc_batch* out);
/* Build a batch straight from models and explicit state lists: instance i is the generator
 * of models[i] (time-invariant) on the n_states[i] states of states[i] (row-major, ns int32
 * each). The state spaces and generators (StateSpaceSparse and FspMatrixSparse of the
 * reference, sparsestatespace.jl:103-144 and fspsparsematrix.jl:47-108) are built on
 * n_workers host threads of the library (0: 4), each on streams of its own, then copied
 * into the batch and freed. Errors name the instance. No reference counterpart (the
 * reference builds one model at a time). */
int nc_batch_create_models(int32_t n, const nc_model* models, const int32_t* const* states, const int64_t* n_states,
                           int32_t precision, int32_t n_workers, nc_batch* out);
int nc_batch_destroy(nc_batch B);
/* Total device rows (padded) and each instance's n_rows (instance_rows: n entries). */
int nc_batch_info(nc_batch B, int64_t* total_rows, int64_t* instance_rows);
/* Integrate every instance from u0 (the instances' n_rows vectors, concatenated) through
 * the shared ascending output times t_out[0..n_out) (t_out[0] is the start); opts gives
 * rtol, atol, h0 and no_graphs. Snapshots stay on the device. */
int nc_batch_solve(nc_batch B, const nc_solve_opts* opts, const double* u0, int32_t n_out,
                   const double* t_out, nc_solve_stats* st);
/* Autotuner of the batched kernels' launch shape: times reps batched DP5(4) iterations at
 * 1, 2 and 4 rows per thread (256, 512 and 1024 rows per workgroup) and keeps the fastest
 * for later solves. Instances are padded to 256-row granules, so every shape runs on the
 * same layout. Outputs (either may be null): the rows per thread chosen and the time of one
 * iteration in ms. No reference counterpart. */
int nc_batch_autotune(nc_batch B, int32_t reps, int32_t* rows_per_thread, double* best_ms);
/* Set the launch shape by hand (1, 2 or 4 rows per thread). */
int nc_batch_set_rows_per_thread(nc_batch B, int32_t rows);
/* Copy snapshot k of instance inst (its n_rows doubles). */
int nc_batch_get_snapshot(nc_batch B, int32_t inst, int32_t k, double* u);
/* Copy snapshot k of every instance in one download, instance after instance (each its
 * n_rows doubles, the layout of nc_batch_solve's u0). No reference counterpart. */
int nc_batch_get_snapshots(nc_batch B, int32_t k, double* u);

/* ======================================================================================
 * Adaptive FSP — replaces solve(model, p0, tspan, ::AdaptiveFspSparse; saveat, fsptol,
 * odeatol, odertol) (fspsolve.jl:59-62,105-197) and the space adapters RStepAdapter /
 * SelectiveRStepAdapter with their init! / adapt! (rstepadapters.jl:12-110).
 * ==================================================================================== */
/* NC_ADAPTER_LEARNED has no reference counterpart: it is SelectiveRStepAdapter whose
 * first expansion level starts only from the boundary states a trained nc_mlp selects
 * (see the learned space adapter section below). */
enum { NC_ADAPTER_RSTEP = 0, NC_ADAPTER_SELECTIVE_RSTEP = 1, NC_ADAPTER_LEARNED = 2 };
typedef struct nc_mlp_s* nc_mlp;
typedef struct {
  int32_t kind;               /* NC_ADAPTER_* (rstepadapters.jl:12-16, 61-65)          */
  int32_t initial_step_count; /* init! expansion depth                                 */
  int32_t max_step_count;     /* adapt! expansion depth                                */
  int32_t dropstates;         /* drop low-probability states before expanding         */
  nc_mlp mlp;                 /* NC_ADAPTER_LEARNED: the trained network (borrowed)    */
  double threshold;           /* NC_ADAPTER_LEARNED: expand from a boundary state when
                                 sigmoid(logit) >= threshold                          */
  double fallback_dt;         /* NC_ADAPTER_LEARNED: when the sink bound fires again
                                 sooner than this after a learned adaptation, the next
                                 adaptation is plain SelectiveRStep; <= 0 means
                                 1e-3 (t1 - t0)                                       */
} nc_adapter;

/* init!(space, adapter, p, t, fsptol) — rstepadapters.jl:23-28, 74-79: expand! over all
 * reactions by initial_step_count levels. The caller appends zeros to its p for the new
 * states (nc_space_count gives the new size). */
int nc_adapter_init(nc_space space, const nc_adapter* ad);

/* adapt!(space, adapter, p, sinks, t, tend, fsptol; integrator) — rstepadapters.jl:35-52,
 * 86-110. p: the n probabilities of the space's states. dsinks: the nr sink entries of
 * du = A(t) u at time t (get_du!, :101-104), required by the selective adapter and
 * ignored otherwise. On return drop_ids[0..*n_drop) holds the sorted 0-based states that
 * were deleted (capacity n); the caller deletes them from p and appends zeros for the
 * states that expand! added. Errors: NC_ERR_ARG for an empty p with the selective adapter
 * (the reference's ArgumentError, :90-92), and for NC_ADAPTER_LEARNED, whose features need
 * the model: it adapts inside nc_solve_adaptive. */
int nc_adapter_adapt(nc_space space, const nc_adapter* ad, int64_t n, const double* p, const double* dsinks,
                     double t, double tend, double fsptol, int64_t* n_drop, int64_t* drop_ids);

typedef struct nc_records_s* nc_records; /* learned-adapter training records (below) */
typedef struct {
  nc_adapter adapter;         /* space_adapter of AdaptiveFspSparse (fspsolve.jl:59-62)  */
  int32_t precision;          /* generator precision, NC_FP64 or NC_FP32                 */
  int32_t max_adaptations;    /* error out after this many adaptations (0 = unlimited)   */
  double fsptol;              /* FSP tolerance (default 1e-6, fspsolve.jl:108)           */
  nc_solve_opts ode;          /* ode_method and its tolerances                           */
  int32_t verbose;            /* print each space update to stdout (fspsolve.jl:184-186) */
  nc_records records;         /* optional: collect learned-adapter training records      */
} nc_adaptive_opts;

/* FspOutputSparse (fspoutput.jl:19-74): snapshots with their own state lists, kept on
 * the device until copied out. */
typedef struct nc_output_s* nc_output;
int nc_adaptive_opts_default(nc_adaptive_opts* opts);
/* Adaptive solve from the initial distribution (n0 states, n0 x ns row-major, unique and
 * non-negative, with probabilities p0) over tspan [t0, t1]. saveat[0..n_save) ascending;
 * n_save == 0 saves the start of every space and the end of every integrator step, as
 * DifferentialEquations does for an empty saveat. The final snapshot appears twice, as
 * in the reference (fspsolve.jl:163-167 and :189-194). */
int nc_solve_adaptive(nc_model model, int64_t n0, const int32_t* states0, const double* p0, double t0, double t1,
                      int32_t n_save, const double* saveat, const nc_adaptive_opts* opts, nc_output* out,
                      nc_solve_stats* st);
/* Adaptive sensitivity solve — solve(model::CmeModelWithSensitivity, initial_condition,
 * tspan, ::AdaptiveForwardSensFspSparse; ...) (forwardsenscmesparse.jl:99-215) with the
 * ForwardSensRStepAdapter (fsspaceadapterssparse.jl:10-60; opts->adapter.kind must be
 * NC_ADAPTER_RSTEP). S0: n_params x n0 initial sensitivities, row-major. The model needs
 * its gradients. Snapshots then also carry S and dsinks (nc_output_get_sens). */
int nc_solve_sens_adaptive(nc_model model, int64_t n0, const int32_t* states0, const double* p0,
                           const double* S0, double t0, double t1, int32_t n_save, const double* saveat,
                           const nc_adaptive_opts* opts, nc_output* out, nc_solve_stats* st);
/* Many independent adaptive solves at once (a parameter sweep). No reference counterpart:
 * the reference solves one CmeModel per solve call, so a sweep is a loop over
 * fspsolve.jl:105-197; each instance here gives exactly what nc_solve_adaptive gives for
 * it. Instance i solves models[i] from its initial distribution: n0[i] states, which follow
 * those of instances 0..i-1 in states0 (n0[i] x ns_i row-major) and p0; every instance
 * shares tspan, saveat and opts (opts->records must be null). n_workers host threads
 * (<= 0: 4) each run whole solves on streams of their own, so that one solve's small
 * launches, host decisions and synchronisations overlap the others' device work; they
 * launch kernels directly (opts->ode.no_graphs is taken as 1), since a stream capture in
 * one thread does not survive the allocations of the others. An instance that runs out of
 * device memory next to the others is solved again, alone, after them. On success outs[i]
 * (and stats[i], if stats is not null) belong to instance i; on an error no output is
 * returned and the message names the first instance that failed. */
int nc_solve_adaptive_batch(int32_t n_inst, const nc_model* models, const int64_t* n0, const int32_t* states0,
                            const double* p0, double t0, double t1, int32_t n_save, const double* saveat,
                            const nc_adaptive_opts* opts, int32_t n_workers, nc_output* outs, nc_solve_stats* stats);
int nc_output_destroy(nc_output o);
/* Number of snapshots, and of state spaces (adaptations + 1). */
int nc_output_count(nc_output o, int32_t* n_snapshots, int32_t* n_spaces);
/* Time, state count and state-space index of snapshot k. */
int nc_output_info(nc_output o, int32_t k, double* t, int64_t* n_states, int32_t* space_id);
/* Copy snapshot k: states (n_states x ns int32), p (n_states) and sinks (nr); any output
 * may be null. */
int nc_output_get(nc_output o, int32_t k, int32_t* states, double* p, double* sinks);
/* Copy the sensitivity of snapshot k to parameter `param` (0-based): S (n_states) and
 * dsinks (nr); either may be null. */
int nc_output_get_sens(nc_output o, int32_t k, int32_t param, double* S, double* dsinks);

/* ======================================================================================
 * Learned space adapter network — no reference counterpart. An 8 -> 64 -> 64 -> 1 MLP
 * (tanh-form GELU hidden layers, keep/add logit) on BF16 MFMA with FP32 accumulation,
 * trained on the device (binary cross-entropy, Adam on FP32 master weights). Parameters
 * (4801 floats): W1 [8][64], b1 [64], W2 [64][64], b2 [64], w3 [64], b3.
 * ==================================================================================== */
#define NC_MLP_FEATURES 8
#define NC_MLP_PARAMS 4801
int nc_mlp_create(uint64_t seed, nc_mlp* out);
int nc_mlp_destroy(nc_mlp M);
int nc_mlp_get_params(nc_mlp M, float* params);
int nc_mlp_set_params(nc_mlp M, const float* params);
/* logits[i] for the n x 8 feature rows x */
int nc_mlp_forward(nc_mlp M, int64_t n, const float* x, float* logits);
/* Gradient of the mean binary cross-entropy over (x, y) (labels 0/1), and the loss. */
int nc_mlp_grad(nc_mlp M, int64_t n, const float* x, const float* y, float* grad, double* loss);
/* Adam training over shuffled mini-batches; *final_loss is the mean loss of the last epoch. */
int nc_mlp_train(nc_mlp M, int64_t n, const float* x, const float* y, int32_t epochs, int32_t batch, float lr,
                 uint64_t seed, double* final_loss);

/* Average duration (ms) of one launch of the forward (train = 0) or forward+backward
 * (train = 1) kernel over n device-resident synthetic feature rows, timed with HIP events
 * around reps back-to-back launches (the MLP's roofline measurement). */
int nc_mlp_bench(nc_mlp M, int64_t n, int32_t reps, int32_t train, double* ms_per_launch);

/* Learned space adapter — no reference counterpart. At a sink-bound event the reactions
 * to expand are chosen as SelectiveRStepAdapter chooses them (rstepadapters.jl:100-108);
 * the boundary states are the states with a sink on one of them. Each boundary state x of
 * the space (n states, probabilities p) at time t gets eight features:
 *   f0..f2  x_s / (1 + max_s) for species s = 0, 1, 2 (max over the space; 0 for s >= ns)
 *   f3      log10(max(p(x), 1e-30)) / 30
 *   f4      log10(1 + sum_r a_r(t, x)) / 4                       (total propensity)
 *   f5      log10(max(share, 1e-30)) / 30, share = p(x) sum_{r in sinks(x)} a_r(t, x) / sum_r dsinks_r
 *   f6      t / tend
 *   f7      |sinks(x)| / nr                                         (all reactions)
 * The network's logit keeps a boundary state when sigmoid(logit) >= threshold; expand!
 * then starts its first level from the kept boundary states only. Nothing kept, or a sink
 * event sooner than fallback_dt after the previous learned adaptation, makes the
 * adaptation plain SelectiveRStep; the sink bound still judges every step. */
/* Boundary states (0-based ids, ascending; capacity n) and their feature rows (n x 8) for
 * the given reactions-to-expand choice (dsinks, nr values). */
int nc_adapter_features(nc_space space, nc_model model, int64_t n, const double* p, const double* dsinks, double t,
                        double tend, int64_t* n_boundary, int64_t* ids, float* features);
/* Training records: during an adaptive solve whose opts->records is set, every adaptation
 * records the features of up to per_adaptation boundary states (an even stride over them),
 * and labels each one at the next adaptation (or at the end of the solve): 1 when the
 * probability then held by its successors along its sink reactions is >= label_mass. */
int nc_records_create(int64_t capacity, int64_t per_adaptation, double label_mass, nc_records* out);
int nc_records_destroy(nc_records R);
int nc_records_count(nc_records R, int64_t* n);
int nc_records_get(nc_records R, float* features, float* labels);
int nc_records_clear(nc_records R);

/* ======================================================================================
 * Sparse FSP vectors on the device — replace FspVectorSparse (src/fspvector/fspvector.jl).
 * A vector holds n states (n x ns int32, row-major) and their n double values. Snapshots
 * of a solve are handed out as vectors without leaving the device; marginals and dense
 * conversion run on the device.
 * ==================================================================================== */
typedef struct nc_fspvec_s* nc_fspvec;
/* FspVectorSparse(states, values) — fspvector.jl:28-35 (the caller checks equal lengths,
 * :29). The states are kept in the given order, duplicates included. */
int nc_fspvec_create(int32_t ns, int64_t n, const int32_t* states, const double* values, nc_fspvec* out);
int nc_fspvec_destroy(nc_fspvec v);
/* nnz(p) — fspvector.jl:24 — and the species count. */
int nc_fspvec_size(nc_fspvec v, int64_t* n, int32_t* ns);
/* get_states / get_values — fspvector.jl:20-21. states: n x ns; either may be null. */
int nc_fspvec_get(nc_fspvec v, int32_t* states, double* values);
/* sum(p) — fspvector.jl:57-59 (a fixed-order tree sum). */
int nc_fspvec_sum(nc_fspvec v, double* total);
/* sum(p, dims) — fspvector.jl:66-99. dims: n_dims 0-based species (repeats allowed) summed
 * out; the result lists the reduced states in order of first appearance, each value the
 * left-to-right sum of its states' values in state order (the reference's loop, bit for
 * bit). NC_ERR_ARG for a dimension outside 0..ns-1 (the reference's ArgumentError). */
int nc_fspvec_marginal(nc_fspvec v, int32_t n_dims, const int32_t* dims, nc_fspvec* out);
/* Array(p) — fspvector.jl:107-129. bounds (ns entries) receives max count + 1 per species;
 * with out non-null (capacity >= prod(bounds) doubles) the dense array is written
 * column-major (species 0 fastest, Julia's layout), zeros elsewhere; for a repeated state
 * the last one wins. NC_ERR_ARG for an empty vector (the reference's ArgumentError). */
int nc_fspvec_dense(nc_fspvec v, int64_t* bounds, int64_t capacity, double* out);
/* Snapshot k of a fixed-space solve as a device vector over the space's states: block 0 is
 * p, block b >= 1 the sensitivity to parameter b-1 of a sensitivity solver (the p /
 * S_b fields of FspOutputSparse, fspoutput.jl:19-23, and sensoutputsparse.jl:20-68).
 * space: the space the solver's generator was built from. sinks (nr doubles) may be null. */
int nc_solver_get_fspvec(nc_solver s, int32_t k, int32_t block, nc_space space, nc_fspvec* p, double* sinks);
/* Snapshot k of an adaptive solve as a device vector: param == -1 gives p, param >= 0 the
 * sensitivity to that parameter. sinks (nr doubles) may be null. */
int nc_output_get_fspvec(nc_output o, int32_t k, int32_t param, nc_fspvec* p, double* sinks);

/* ---- helpers exposed for testing --------------------------------------------------- */
/* Dense matrix exponential of an n x n row-major matrix (Pade 13 scaling and squaring,
 * Higham 2005) — the host-side Hessenberg exponential of the Krylov integrator. */
int nc_expm_dense(int32_t n, const double* a, double* out);

#ifdef __cplusplus
}
#endif
#endif /* NUMCME_HIP_H */
