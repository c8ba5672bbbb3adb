// FSP generator storage and mat-vec (internal).
//
// Layout in HBM for a space of n states and nr reactions (VT = float or double):
//   diag    double [ld]      merged time-invariant diagonal
//   col     int32  [K][ld]   merged time-invariant off-diagonal columns, slot-major
//   val     VT     [K][ld]   (rows with fewer than K predecessors are padded: col = row, val = 0)
//   tvd     VT     [T][ld]   per time-varying reaction: diagonal (statefactor or joint value)
//   tvc     int32  [T][ld]   per time-varying reaction: predecessor column (or row)
//   tvv     VT     [T][ld]   per time-varying reaction: predecessor value (or 0)
//   (ld = n rounded up to 64; rows n..ld-1 hold zeros and column 0)
//   sink_off int64 [nr+1]    sink row r holds entries sink_off[r] .. sink_off[r+1)
//   sink_src int32 [E]       source state of each sink entry
//   sink_val VT    [E]
// Row i of A(t) = diag*v_i + sum_k val*v[col] + sum_q scale_q(t)*(tvd*v_i + tvv*v[tvc]);
// sink row r = scale_r(t) * sum_e sink_val*v[sink_src]. Slot-major arrays make every
// load of a wavefront one contiguous 256-byte (FP32) or 512-byte (FP64) segment.
#pragma once

#include "nc_model.h"
#include "nc_space.h"

struct nc_matrix_s {
  int precision = NC_FP64;
  size_t vsize = 8;  // bytes per stored value / vector entry
  int ns = 0, nr = 0;
  int64_t n = 0;       // states
  int64_t n_rows = 0;  // states + sinks
  int64_t ld = 0;      // slot stride (n rounded up to 64)
  int K = 0;
  int rows_per_thread = 2;  // mat-vec rows per thread (1, 2 or 4; see nc_matrix_autotune)
  uint64_t launch_gen = 0;  // bumped whenever the launch shape changes: graphs captured
                            // with an older shape are re-captured by their owners
  nc::DevBuf<double> diag;
  nc::DevBuf<int32_t> col;
  nc::DevBuf<unsigned char> val;
  std::vector<int> tv_r;     // reaction index of each time-varying reaction
  std::vector<int> tv_kind;  // NC_PROP_SEPARABLE or NC_PROP_JOINT
  nc::DevBuf<unsigned char> tvd, tvv;
  nc::DevBuf<int32_t> tvc;
  std::vector<int64_t> sink_off_h;
  nc::DevBuf<int64_t> sink_off;
  nc::DevBuf<int32_t> sink_src;
  nc::DevBuf<unsigned char> sink_val;
  // long sink rows (adaptive spaces have large boundaries) are reduced in fixed chunks by
  // many workgroups and folded in order by a second kernel; 0 chunks: one workgroup per row
  int32_t n_chunks = 0;
  nc::DevBuf<int64_t> chunk_start;  // n_chunks + 1 entry offsets
  nc::DevBuf<int32_t> chunk_row;    // sink row of each chunk
  nc::DevBuf<int32_t> chunk_first;  // nr + 1: first chunk of each sink row
  nc::DevBuf<double> chunk_part;    // n_chunks partial sums
  nc::DevBuf<unsigned> chunk_arrivals;  // chunk workgroups done in the running launch (a
                                        // matrix runs one mat-vec at a time)
  // slots whose stored column is i + dia_off[k] in every row with a nonzero value (box
  // spaces in lattice order): the mat-vec computes those columns instead of loading them
  uint32_t dia_mask = 0;
  std::vector<int32_t> dia_off;
  int64_t max_reach = 0;  // largest |column - row| over the slots' nonzero entries
  // Computed mat-vecs: on box spaces in lattice order (species 0 fastest) whose
  // propensities are all time-invariant and each read at most one species, every value is
  // a lookup in its reaction's table (the propensity over its species' range, evaluated by
  // the build's own program evaluator) and every column a row offset; mode 1 still reads
  // the FP64 diagonal, mode 2 computes it too and reads only v (and the sink lists).
  bool mf_ok = false;     // the build found the space and model eligible
  bool mf_multi = false;  // some slot holds several reactions or changes several species
  int mf_mode = 0;        // 0: stored arrays; 1: computed values, stored diagonal (needs
                          // !mf_multi); 2: matrix-free (values and diagonal computed)
  nc::DevBuf<double> mf_tab;            // concatenated per-reaction tables
  nc::DevBuf<unsigned char> mf_info;    // box, groups and table offsets (MfInfo)
  int64_t mf_table_entries = 0;
  nc::DevBuf<int32_t> states;      // kept for joint refreshes
  nc::DevBuf<int32_t> state_conn;  // kept for joint refreshes
  nc_model_s* model = nullptr;     // borrowed; must outlive the matrix
  double t_cache = -INFINITY;
  nc::Stream stream;
  // reporting
  int64_t nnz_ref = 0, n_sink_entries = 0;
  double norm1 = 0.0;
  // kernel timing (HIP events around mat-vec launches); timing_plain: the launches timed
  // are plain mat-vecs (the Krylov projection sums then run as a separate pass)
  bool timing = false;
  bool timing_plain = false;
  std::vector<hipEvent_t> ev;
  size_t ev_used = 0;
  double timed_ms = 0.0;
  int64_t timed_count = 0;
  int64_t launches = 0;
  ~nc_matrix_s();
};

namespace nc {
// Bytes of the fixed device scratch behind nc_matrix_matvec_noalloc (per device).
constexpr size_t kNoAllocScratchBytes = size_t(48) << 20;
// FspMatrixSparse{RealT}(space, propensities; parameters) for the space's current states
nc_matrix_s* matrix_create(nc_space_s* sp, nc_model_s* m, int precision);
// Bring time-dependent values up to date for time t (joint refresh, on the matrix stream).
void matrix_prepare(nc_matrix_s* A, double t);
// Launch-time scales at t: sink rows (nr) and time-varying slots (tv_r.size()).
void matrix_scales(const nc_matrix_s* A, double t, double* sink_scale, double* tv_scale);
// out = A(t) v (+ out when accumulate); v is a device vector of n_rows VT values, out holds
// n_rows VT values, or FP64 values when out_fp64 is set.
void matvec_launch(nc_matrix_s* A, double t, const void* v, void* out, bool accumulate, hipStream_t s,
                   bool out_fp64 = false);
// A DP5(4) stage input formed in the mat-vec's epilogue from the stage vector the launch
// produces: y = u + h (sum_{j: bit j of mask} a[j] k[j] + a_new k_new), written as the
// next mat-vec's input copy vin (VT, a buffer other than the one the launch reads) and, when
// ytmp is set, in FP64 there (the last stage, y_{n+1}). State and sink rows are written;
// padding rows are not.
struct RkEpilogue {
  const double* k[6];
  double a[6];
  double a_new;
  int mask;
  double h;
  const double* u;
  double* ytmp;
  void* vin;
};
// Whether the matrix's mat-vec kernel takes the epilogue (the stored generator at 2 rows per
// thread); otherwise the stage pass runs on its own.
bool matvec_rk_fusable(const nc_matrix_s* A);
// out = A(t) v (FP64) and the epilogue's stage input
void matvec_launch_rk(nc_matrix_s* A, double t, const void* v, double* out, hipStream_t s, const RkEpilogue& epi);
// The DP5(4) error pass in the epilogue of the step's last mat-vec, the one that forms
// k_7 = f(y_{n+1}) (the same generators as RkEpilogue): per row
// r^2 = (h sum_j e_j k_j / (atol + rtol max(|u|, |y_{n+1}|)))^2, summed per row workgroup
// into part[b] (the mat-vec's row workgroups in row order) and per sink row r into
// part[blocks + r], blocks = matvec_err_blocks(A) - nr; with n_ex > 0 also the sink values
// the event check reads: ex[r] = y_{n+1}, ex[nr + r] = k_1, ex[2 nr + r] = k_7 at sink row r.
struct RkErrEpilogue {
  const double* k[6];  // k_1 .. k_6 (k_2 is not read: its weight is zero)
  double e[7];         // the error weights; e[6] multiplies the new k_7
  double h, atol, rtol;
  const double* u;
  const double* ynew;
  double* part;  // device-visible (the solver passes pinned host memory)
  double* ex;
  int n_ex;
};
int64_t matvec_err_parts(const nc_matrix_s* A);  // row workgroups + sink rows
void matvec_launch_err(nc_matrix_s* A, double t, const void* v, double* out, hipStream_t s, const RkErrEpilogue& epi);
// Arnoldi projection sums computed by the mat-vec (nc_dual.h): besides out = A v, every
// row workgroup writes the sums of its 512 rows, with v = V_{nv-1} and q = A v, into
// part (value-major, nblk = the mat-vec's row workgroups). Time-invariant stored
// generators with 2 rows per thread and 1..8 slots whose columns are all loaded or all
// computed, unless plain mat-vecs are being timed; other generators take the separate
// block pass.
struct DualDots {
  const void* V;     // V_0 .. V_{nv-1}, matrix precision, stride ld
  int64_t ld;
  int32_t nv;        // 1..32
  const double* sc;  // scale of each V_i . q sum (null: 1)
  double* part;
  int64_t nblk;
};
bool matvec_dots_fusable(const nc_matrix_s* A);
void matvec_launch_dots(nc_matrix_s* A, const void* v, double* out, hipStream_t s, const DualDots& d);
// When kernel timing is on, hand out the next pair of HIP events for a dispatch-stamped
// launch (the caller increments ev_used after the launch); false when timing is off.
bool matvec_timing_slot(nc_matrix_s* A, hipEvent_t* e0, hipEvent_t* e1);
// Fold recorded kernel timings into timed_ms (synchronises on the recorded events).
void matvec_harvest_timing(nc_matrix_s* A);
double matrix_bytes_per_matvec(const nc_matrix_s* A);
}  // namespace nc
