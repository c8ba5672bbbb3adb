// FSP generator build and mat-vec on the device.
//
// Replaces FspMatrixSparse (NumCME.jl src/fspmatrix/sparse/fspsparsematrix.jl):
//  * the constructor (:47-108) and _generate_sparsematrix_entries (:120-152) become one
//    count kernel and one fill kernel over the state rows plus a sink-list build;
//    time-invariant reactions with identical stoichiometry share a predecessor and are
//    merged into one slot, exactly the duplicates SparseArrays.sparse sums (:74);
//  * _update_sparsematrix! (:154-166) becomes a refresh kernel for joint time-varying
//    reactions, run when t differs from the cached time (:206-213);
//  * matvec!/matvecadd! (:196-247) become one kernel: R consecutive state rows per
//    thread (slot-major ELL, 8/16-byte vector loads of the matrix streams, FP64
//    accumulation) and one workgroup per sink row (segmented reduction). R is chosen by
//    nc_matrix_autotune.
#include <algorithm>
#include <cfloat>

#include <hip/hip_ext.h>

#include <mutex>

#include "nc_dual.h"
#include "nc_matrix.h"

nc_matrix_s::~nc_matrix_s() {
  for (auto e : ev) (void)hipEventDestroy(e);
}

namespace nc {

namespace {

constexpr int kBlock = 256;
constexpr int64_t kLdAlign = 64;  // slot stride in elements: keeps every slot 256-byte aligned

struct Groups {
  int32_t ng;
  int32_t leader[kMaxReactions];   // first reaction of each group
  int32_t start[kMaxReactions + 1];
  int32_t members[kMaxReactions];  // reactions, grouped, ascending within a group
};

struct TvList {
  int32_t n;
  int32_t r[kMaxReactions];
  int32_t kind[kMaxReactions];
};

struct Progs {
  const nc_instr* code;
  const int64_t* offs;  // 2*nr+1
  const double* theta;
  int32_t reg;          // every program fits prog_eval_reg's stack
  int32_t ns;
};

__device__ inline double xprog(const Progs& p, int r, double t, const int32_t* x) {
  const int64_t o = p.offs[2 * r];
  return prog_eval(p.code + o, p.offs[2 * r + 1] - o, t, x, p.theta);
}

// xprog with the species values from xs(index): the register evaluator when every program
// fits it, else the array evaluator on a copy of the state
template <bool REG, class XS>
__device__ inline double xprog_f(const Progs& p, int r, double t, XS xs) {
  const int64_t o = p.offs[2 * r];
  if constexpr (REG) {
    return prog_eval_reg(p.code + o, p.offs[2 * r + 1] - o, t, xs, p.theta);
  } else {
    int32_t x[kMaxSpecies];
    for (int sx = 0; sx < p.ns; ++sx) x[sx] = xs(sx);
    return prog_eval(p.code + o, p.offs[2 * r + 1] - o, t, x, p.theta);
  }
}

Progs make_progs(const nc_model_s* m) {
  return Progs{m->d_code.p, m->d_offs.p, m->d_params.p, m->max_depth <= kRegStack ? 1 : 0, m->ns};
}

// Matrix-free mat-vec metadata (device memory; every kernel read is wave-uniform): the box
// [lo, hi] in lattice order with species 0 fastest, the reaction groups (one slot each, in
// slot order), each group leader's stoichiometry and column offset, the species each
// reaction's propensity reads (-1: none) and where its table starts.
constexpr int kMfSpecies = 4;
constexpr int kMfSlots = 8;
constexpr int kMfMembers = 4;  // reactions per group (one stoichiometry) in the matrix-free mode
struct MfInfo {
  int32_t ns, ng;
  int32_t lo[kMfSpecies], hi[kMfSpecies];
  int64_t stride[kMfSpecies];
  int32_t start[kMfSlots + 1];
  int32_t members[kMaxReactions];
  int32_t nu[kMfSlots][kMfSpecies];
  int64_t off[kMfSlots];  // column of the slot's predecessor: i - off
  int32_t sp[kMaxReactions];
  int64_t toff[kMaxReactions];
  int32_t cs[kMfSlots];  // the one species the slot's stoichiometry changes (lean mode)
};

__global__ void k_build_count(int64_t n, int nr, const int32_t* state_conn, Groups g, TvList tv, int32_t* kmax,
                              unsigned long long* offdiag_total, unsigned long long* tv_pred) {
  // Grid-stride over the states with a bounded grid, reduced over the workgroup: one atomic
  // per workgroup and counter (same-address atomics from every wave serialised this
  // kernel). The first eight groups' connectivity loads of a row go out together.
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int wmax = 0;
  unsigned long long wsum = 0;
  unsigned long long tvn[4] = {0, 0, 0, 0};  // time-varying slots 0..3 (the rest per row below)
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    int cnt = 0;
    if (g.ng > 0) {
      const int64_t row = i * nr;
      int32_t c[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) c[k] = state_conn[row + g.leader[min(k, g.ng - 1)]];
#pragma unroll
      for (int k = 0; k < 8; ++k) cnt += (k < g.ng && c[k] >= 0) ? 1 : 0;
      for (int k = 8; k < g.ng; ++k) cnt += state_conn[row + g.leader[k]] >= 0;
    }
    wmax = max(wmax, cnt);
    wsum += (unsigned long long)cnt;
    for (int q = 0; q < tv.n; ++q) {
      const bool has = state_conn[i * nr + tv.r[q]] >= 0;
      if (q < 4)
        tvn[q] += has ? 1ull : 0ull;
      else if (has)
        atomicAdd(&tv_pred[q], 1ull);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    wmax = max(wmax, __shfl_down(wmax, off, 64));
    wsum += __shfl_down(wsum, off, 64);
#pragma unroll
    for (int q = 0; q < 4; ++q) tvn[q] += __shfl_down(tvn[q], off, 64);
  }
  __shared__ int s_max[kBlock / 64];
  __shared__ unsigned long long s_sum[kBlock / 64], s_tv[4][kBlock / 64];
  if (lane == 0) {
    s_max[wave] = wmax;
    s_sum[wave] = wsum;
#pragma unroll
    for (int q = 0; q < 4; ++q) s_tv[q][wave] = tvn[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int bmax = 0;
    unsigned long long bsum = 0, btv[4] = {0, 0, 0, 0};
    for (int w = 0; w < kBlock / 64; ++w) {
      bmax = max(bmax, s_max[w]);
      bsum += s_sum[w];
      for (int q = 0; q < 4; ++q) btv[q] += s_tv[q][w];
    }
    atomicMax(kmax, bmax);
    if (bsum) atomicAdd(offdiag_total, bsum);
    for (int q = 0; q < 4 && q < tv.n; ++q)
      if (btv[q]) atomicAdd(&tv_pred[q], btv[q]);
  }
}

// Rows i in [n, ld) are padding: zero values and column 0, so vectorised reads of a
// thread's row group never leave the arrays and never gather out of range.
// With `positional` (set when some row has every group's predecessor, so K equals the
// group count) group k always occupies slot k and a missing predecessor is a zero entry
// in column i; the products then run in the same order with zeros added, so the row sums
// are unchanged, and on lattice-ordered spaces each slot's column is row + constant.
// Otherwise a row's entries are packed into its first slots.
// One row of the build; returns |diag| (0 for a padding row).
template <typename VT, bool REG>
__device__ double build_fill_row(int64_t i, int64_t n, int64_t ld, int ns, int nr, int K, int positional,
                                 const int32_t* states, const int32_t* state_conn, const uint32_t* sink_mask,
                                 const int32_t* stoich, const Groups& g, const TvList& tv, const Progs& pr,
                                 double* diag, int32_t* col, VT* val, VT* tvd, int32_t* tvc, VT* tvv) {
  if (i >= n) {
    diag[i] = 0.0;
    for (int k = 0; k < K; ++k) col[k * ld + i] = 0, val[k * ld + i] = (VT)0;
    for (int q = 0; q < tv.n; ++q) tvd[q * ld + i] = (VT)0, tvv[q * ld + i] = (VT)0, tvc[q * ld + i] = 0;
    return 0.0;
  }
  const int32_t* xrow = states + i * ns;
  auto xs = [&](int a) { return xrow[a]; };
  const uint32_t smask = sink_mask[i];
  int slot = 0;
  double d = 0.0;
  for (int k = 0; k < g.ng; ++k) {
    const int lead = g.leader[k];
    // incoming entry of this row from the group's predecessor x - s_g
    const int32_t j = state_conn[i * nr + lead];
    if (j >= 0) {
      auto ys = [&](int a) { return xrow[a] - stoich[a + lead * ns]; };
      double a = 0.0;
      for (int m = g.start[k]; m < g.start[k + 1]; ++m) a += xprog_f<REG>(pr, g.members[m], 0.0, ys);
      col[slot * ld + i] = j;
      val[slot * ld + i] = (VT)a;
      ++slot;
    } else if (positional) {
      col[slot * ld + i] = (int32_t)i;
      val[slot * ld + i] = (VT)0;
      ++slot;
    }
    // this row's outflow through the group: the diagonal is minus the stored column
    // entries, so each stored column sums to zero exactly in real arithmetic
    bool nonneg = true;
    for (int s = 0; s < ns; ++s) nonneg = nonneg && (xrow[s] + stoich[s + lead * ns] >= 0);
    if (!nonneg) {
      // no successor and no sink: the mass leaks, as in the reference (:129-132)
      for (int m = g.start[k]; m < g.start[k + 1]; ++m) d += xprog_f<REG>(pr, g.members[m], 0.0, xs);
    } else if (smask & (1u << lead)) {
      for (int m = g.start[k]; m < g.start[k + 1]; ++m) d += (double)(VT)xprog_f<REG>(pr, g.members[m], 0.0, xs);
    } else {
      double a = 0.0;
      for (int m = g.start[k]; m < g.start[k + 1]; ++m) a += xprog_f<REG>(pr, g.members[m], 0.0, xs);
      d += (double)(VT)a;
    }
  }
  for (; slot < K; ++slot) {
    col[slot * ld + i] = (int32_t)i;
    val[slot * ld + i] = (VT)0;
  }
  diag[i] = -d;
  const double dabs = d > 0 ? d : -d;
  for (int q = 0; q < tv.n; ++q) {
    const int r = tv.r[q];
    const int32_t j = state_conn[i * nr + r];
    if (tv.kind[q] == NC_PROP_SEPARABLE) {
      tvd[q * ld + i] = -(VT)xprog_f<REG>(pr, r, 0.0, xs);
      if (j >= 0) {
        auto ys = [&](int a) { return xrow[a] - stoich[a + r * ns]; };
        tvv[q * ld + i] = (VT)xprog_f<REG>(pr, r, 0.0, ys);
      } else {
        tvv[q * ld + i] = (VT)0;
      }
    } else {  // joint: values are written by the refresh kernel
      tvd[q * ld + i] = (VT)0;
      tvv[q * ld + i] = (VT)0;
    }
    tvc[q * ld + i] = j >= 0 ? j : (int32_t)i;
  }
  return dabs;
}

// The build's rows, one per thread, and the largest |diag| (nonnegative doubles order as
// their bit patterns), reduced over the wavefront before its one atomic.
template <typename VT, bool REG>
__global__ __launch_bounds__(kBlock) void k_build_fill(int64_t n, int64_t ld, int ns, int nr, int K, int positional, const int32_t* states,
                             const int32_t* state_conn, const uint32_t* sink_mask, const int32_t* stoich, Groups g,
                             TvList tv, Progs pr, double* diag, int32_t* col, VT* val, VT* tvd, int32_t* tvc,
                             VT* tvv, unsigned long long* diag_max_bits) {
  // one row per thread (the rows' propensity programs keep every wave busy); |diag|
  // reduced over the workgroup: one atomic per workgroup
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  double dabs = 0.0;
  if (i < ld)
    dabs = build_fill_row<VT, REG>(i, n, ld, ns, nr, K, positional, states, state_conn, sink_mask, stoich, g, tv, pr, diag,
                              col, val, tvd, tvc, tvv);
  for (int off = 32; off > 0; off >>= 1) dabs = fmax(dabs, __shfl_down(dabs, off, 64));
  __shared__ double s_d[kBlock / 64];
  if ((threadIdx.x & 63) == 0) s_d[threadIdx.x >> 6] = dabs;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / 64; ++w) dabs = fmax(dabs, s_d[w]);
    atomicMax(diag_max_bits, (unsigned long long)__double_as_longlong(dabs));
  }
}

// Range of col[i] - i over the rows of slot k = blockIdx.y whose value is nonzero:
// mnmx[2k] = min, mnmx[2k + 1] = max (a grid-stride pass taking four rows per step with
// their loads issued together, wave and workgroup reductions, one atomic pair per
// workgroup). Every slot in one launch.
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_dia_detect(int64_t n, int64_t ld, const int32_t* col_all,
                                                        const VT* val_all, int32_t* mnmx_all) {
  constexpr int U = 4;
  const int32_t* col = col_all + (int64_t)blockIdx.y * ld;
  const VT* val = val_all + (int64_t)blockIdx.y * ld;
  int32_t* mnmx = mnmx_all + 2 * blockIdx.y;
  __shared__ int32_t slo[kBlock / 64], shi[kBlock / 64];
  int32_t lo = INT32_MAX, hi = INT32_MIN;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t ib = blockIdx.x * (int64_t)kBlock + threadIdx.x; ib < n; ib += stride * U) {
    VT w[U];
    int32_t c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = ib + u * stride, ic = i < n ? i : n - 1;
      w[u] = val[ic];
      c[u] = col[ic];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = ib + u * stride;
      if (i < n && w[u] != (VT)0) {
        const int32_t d = c[u] - (int32_t)i;
        lo = min(lo, d);
        hi = max(hi, d);
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    lo = min(lo, __shfl_xor(lo, off, 64));
    hi = max(hi, __shfl_xor(hi, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    slo[threadIdx.x >> 6] = lo;
    shi[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / 64; ++w) lo = min(lo, slo[w]), hi = max(hi, shi[w]);
    atomicMin(mnmx, lo);
    atomicMax(mnmx + 1, hi);
  }
}

// out[i] = in[i * stride] for i < n (n <= 64; one wave)
__global__ void k_gather_strided(const uint32_t* in, int64_t stride, int n, uint32_t* out) {
  const int i = threadIdx.x;
  if (i < n) out[i] = in[(int64_t)i * stride];
}


__global__ void k_sink_flags(int64_t n, int nr, const uint32_t* sink_mask, uint32_t* flags) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n * nr) return;
  const int r = (int)(t / n);
  const int64_t j = t % n;
  flags[t] = (sink_mask[j] >> r) & 1u;
}

template <typename VT, bool REG>
__global__ __launch_bounds__(kBlock) void k_sink_fill(int64_t n, int ns, int nr, const int32_t* states, const uint32_t* flags,
                            const uint32_t* pos, const int32_t* kinds, Progs pr, int32_t* sink_src, VT* sink_val) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n * nr || !flags[t]) return;
  const int r = (int)(t / n);
  const int64_t j = t % n;
  const int32_t* xrow = states + j * ns;
  const uint32_t e = pos[t];
  sink_src[e] = (int32_t)j;
  sink_val[e] = kinds[r] == NC_PROP_JOINT ? (VT)0 : (VT)xprog_f<REG>(pr, r, 0.0, [&](int a) { return xrow[a]; });
}

// _update_sparsematrix! (fspsparsematrix.jl:154-166) for every joint reaction at time t.
template <typename VT>
__global__ void k_joint_refresh(int64_t n, int64_t ld, int ns, int nr, const int32_t* states,
                                const int32_t* state_conn, const int32_t* stoich, TvList tv, Progs pr, double t,
                                VT* tvd, VT* tvv) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t x[kMaxSpecies], y[kMaxSpecies];
  for (int s = 0; s < ns; ++s) x[s] = states[i * ns + s];
  for (int q = 0; q < tv.n; ++q) {
    if (tv.kind[q] != NC_PROP_JOINT) continue;
    const int r = tv.r[q];
    tvd[q * ld + i] = -(VT)xprog(pr, r, t, x);
    const int32_t j = state_conn[i * nr + r];
    if (j >= 0) {
      for (int s = 0; s < ns; ++s) y[s] = x[s] - stoich[s + r * ns];
      tvv[q * ld + i] = (VT)xprog(pr, r, t, y);
    }
  }
}

template <typename VT>
__global__ void k_joint_sink_refresh(int64_t e0, int64_t e1, int r, int ns, const int32_t* states,
                                     const int32_t* sink_src, Progs pr, double t, VT* sink_val) {
  int64_t e = e0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= e1) return;
  int32_t x[kMaxSpecies];
  const int64_t j = sink_src[e];
  for (int s = 0; s < ns; ++s) x[s] = states[j * ns + s];
  sink_val[e] = (VT)xprog(pr, r, t, x);
}

// ---- mat-vec ---------------------------------------------------------------------------
template <typename VT>
struct MatArgs {
  int64_t n;
  int64_t ld;
  int32_t K;
  int32_t ntv;
  int32_t nr;
  const double* diag;
  const int32_t* col;
  const VT* val;
  const VT* tvd;
  const int32_t* tvc;
  const VT* tvv;
  const int64_t* sink_off;
  const int32_t* sink_src;
  const VT* sink_val;
  double tv_scale[kMaxReactions];
  double sink_scale[kMaxReactions];
  uint32_t dia_mask;              // bit k: slot k's column is i + dia_off[k]
  int32_t dia_off[kMaxReactions];
  int32_t n_chunks;               // > 0: the extra workgroups reduce sink-entry chunks
  const int64_t* chunk_start;
  const int32_t* chunk_row;
  const int32_t* chunk_first;     // nr + 1: first chunk of each sink row
  double* chunk_part;
  unsigned* chunk_arrivals;       // chunk workgroups done in this launch, zero between launches
  const MfInfo* mf;               // matrix-free mode (DM 3): metadata and tables
  const double* mf_tab;
  // DP5(4) epilogue (EPI kernels): the next stage's input from the stage vector just
  // formed, y = u + h (sum_{j: bit j of rk_mask} rk_a[j] k_j + rk_a[6] k_new)
  const double* rk_k[6];  // k_j (a duplicate of rk_k[0] where the mask bit is clear)
  double rk_a[7];
  double rk_h;
  const double* rk_u;
  double* rk_ytmp;        // the last stage: y_{n+1} in FP64 too (else null)
  VT* rk_vin;             // the next mat-vec's input copy (not the buffer this launch reads)
  int32_t rk_mask;
  // DP5(4) error pass (ERR kernels)
  const double* err_k[6];
  double err_e[7];
  double err_h, err_atol, err_rtol;
  const double* err_u;
  const double* err_ynew;
  double* err_part;
  double* err_ex;
  int32_t err_nex;
  // Arnoldi projection sums (DOT kernels): basis V_0..V_{dot_nv-1} with stride dot_ld
  const VT* dot_V;
  int64_t dot_ld;
  int32_t dot_nv;
  const double* dot_sc;
  double* dot_part;
};

template <typename T, int R>
struct alignas(sizeof(T) * R) Pack {
  T v[R];
};

template <typename T, int R>
__device__ inline Pack<T, R> ldpack(const T* p) {
  return *reinterpret_cast<const Pack<T, R>*>(p);
}

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Computed-value rows (mode 1): the slot values come from the tables and the columns from
// row offsets, as in the matrix-free form below, but the FP64 diagonal is read from the
// stored generator. Every slot holds one reaction whose stoichiometry changes one species
// (cs[k] by nu[k]), so a slot costs one comparison, one index and one table lookup. Same
// arithmetic, and the same result bit for bit, as the stored kernel.
template <typename VT, typename OT, int KT, int R>
__device__ inline void mf_rows_values(const MatArgs<VT>& a, const VT* __restrict__ v, OT* __restrict__ out,
                                      int64_t i0, int accumulate) {
  const MfInfo* mi = a.mf;
  const int ns = mi->ns;
  const uint32_t n32 = (uint32_t)a.n, i032 = (uint32_t)i0;
  int32_t lo[kMfSpecies], hi[kMfSpecies];
#pragma unroll
  for (int sx = 0; sx < kMfSpecies; ++sx) {
    lo[sx] = sx < ns ? mi->lo[sx] : 0;
    hi[sx] = sx < ns ? mi->hi[sx] : 0;
  }
  int32_t x[kMfSpecies];
  {
    uint32_t rem = i032;
#pragma unroll
    for (int sx = 0; sx < kMfSpecies; ++sx) {
      x[sx] = 0;
      if (sx + 1 < ns) {
        const uint32_t e = (uint32_t)(hi[sx] - lo[sx] + 1), qt = rem / e;
        x[sx] = lo[sx] + (int32_t)(rem - qt * e);
        rem = qt;
      } else if (sx + 1 == ns) {
        x[sx] = lo[sx] + (int32_t)rem;
      }
    }
  }
  uint32_t cc[R][KT], iin[R][KT];
  bool in[R][KT];
#pragma unroll
  for (int r = 0; r < R; ++r) {
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      const int cs = mi->cs[k];
      const int m = mi->members[mi->start[k]];
      const int spm = mi->sp[m];
      const int32_t nu = mi->nu[k][cs];
      int32_t xc = x[0], xm = x[0];
#pragma unroll
      for (int sx = 1; sx < kMfSpecies; ++sx) {
        if (cs == sx) xc = x[sx];
        if (spm == sx) xm = x[sx];
      }
      int32_t lc = lo[0], hc = hi[0], lm = lo[0], wm = hi[0] - lo[0];
#pragma unroll
      for (int sx = 1; sx < kMfSpecies; ++sx) {
        if (cs == sx) lc = lo[sx], hc = hi[sx];
        if (spm == sx) lm = lo[sx], wm = hi[sx] - lo[sx];
      }
      const int32_t yc = xc - nu;
      const uint32_t i = i032 + r;
      in[r][k] = yc >= lc && yc <= hc;
      cc[r][k] = in[r][k] ? i - (uint32_t)mi->off[k] : i;
      const int32_t ym = spm == cs ? yc : xm;
      iin[r][k] = spm < 0 ? 0u : (uint32_t)min(max(ym - lm, 0), wm);  // clamped: no predecessor adds nothing
    }
    x[0] += 1;
#pragma unroll
    for (int sx = 0; sx + 1 < kMfSpecies; ++sx)
      if (sx + 1 < ns && x[sx] > hi[sx]) {
        x[sx] = lo[sx];
        x[sx + 1] += 1;
      }
  }
  double tin[R][KT], gv[R][KT], vi[R], dg[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t ic = min(i032 + r, n32 - 1);
    vi[r] = (double)v[ic];
    dg[r] = a.diag[ic];
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      gv[r][k] = (double)v[cc[r][k]];
      tin[r][k] = a.mf_tab[mi->toff[mi->members[mi->start[k]]] + iin[r][k]];
    }
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t i = i032 + r;
    if (i >= n32) continue;
    double acc = __dmul_rn(dg[r], vi[r]);
#pragma unroll
    for (int k = 0; k < KT; ++k) acc = __fma_rn(in[r][k] ? (double)(VT)tin[r][k] : 0.0, gv[r][k], acc);
    out[i] = (OT)(accumulate ? (double)out[i] + acc : acc);
  }
}

// Matrix-free rows [i0, i0 + R) of a box space in lattice order (species 0 fastest), MM
// members per reaction group at most. The coordinates of row i0 come from one division per
// species, those of the next rows from a step of species 0 with carries. Every index of
// every row is computed first, then every table value and every gather is loaded in one
// batch (indices clamped, unused members weighted out), then the arithmetic runs in the
// build's order: the predecessor entry is the group's FP64 sum rounded to VT, the outflow
// is added per member for leaks and sinks and per group otherwise, and the row sum is
// diag * v_i followed by one fused multiply-add per slot, as in the stored kernel; the
// result equals the stored generator's bit for bit.
template <typename VT, typename OT, int KT, int R, int MM>
__device__ inline void mf_rows(const MatArgs<VT>& a, const VT* __restrict__ v, OT* __restrict__ out, int64_t i0,
                               int accumulate) {
  const MfInfo* mi = a.mf;
  const int ns = mi->ns;
  int32_t lo[kMfSpecies], hi[kMfSpecies], nu[KT][kMfSpecies];
  int32_t off[KT];
  int32_t cnt[KT], spm[KT][MM];
  const double* tb[KT][MM];
#pragma unroll
  for (int sx = 0; sx < kMfSpecies; ++sx) {
    lo[sx] = sx < ns ? mi->lo[sx] : 0;
    hi[sx] = sx < ns ? mi->hi[sx] : 0;
  }
#pragma unroll
  for (int k = 0; k < KT; ++k) {
    off[k] = (int32_t)mi->off[k];
#pragma unroll
    for (int sx = 0; sx < kMfSpecies; ++sx) nu[k][sx] = sx < ns ? mi->nu[k][sx] : 0;
    const int q0 = mi->start[k];
    cnt[k] = mi->start[k + 1] - q0;
#pragma unroll
    for (int q = 0; q < MM; ++q) {
      const int m = mi->members[q < cnt[k] ? q0 + q : q0];
      spm[k][q] = mi->sp[m];
      tb[k][q] = a.mf_tab + mi->toff[m];
    }
  }
  // 32-bit index arithmetic: eligible spaces hold fewer than 2^31 states
  const int32_t n32 = (int32_t)a.n, i032 = (int32_t)i0;
  int32_t x[kMfSpecies];
  {
    uint32_t rem = (uint32_t)i032;
#pragma unroll
    for (int sx = 0; sx < kMfSpecies; ++sx) {
      x[sx] = 0;
      if (sx + 1 < ns) {
        const uint32_t e = (uint32_t)(hi[sx] - lo[sx] + 1), qt = rem / e;
        x[sx] = lo[sx] + (int32_t)(rem - qt * e);
        rem = qt;
      } else if (sx + 1 == ns) {
        x[sx] = lo[sx] + (int32_t)rem;  // the slowest species needs no remainder
      }
    }
  }
  // indices and flags of every row
  int32_t cc[R][KT];
  int32_t iin[R][KT][MM], iout[R][KT][MM];
  bool in[R][KT], leak[R][KT], sink[R][KT], live[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int32_t i = i032 + r;
    live[r] = i < n32;
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      bool pin = true, nonneg = true, sin = true;
      int32_t y[kMfSpecies];
#pragma unroll
      for (int sx = 0; sx < kMfSpecies; ++sx) {
        y[sx] = x[sx];
        if (sx < ns && nu[k][sx] != 0) {
          y[sx] = x[sx] - nu[k][sx];
          const int32_t z = x[sx] + nu[k][sx];
          pin = pin && y[sx] >= lo[sx] && y[sx] <= hi[sx];
          nonneg = nonneg && z >= 0;
          sin = sin && z >= lo[sx] && z <= hi[sx];
        }
      }
      in[r][k] = pin;
      leak[r][k] = !nonneg;
      sink[r][k] = nonneg && !sin;
      cc[r][k] = pin ? i - off[k] : i;
#pragma unroll
      for (int q = 0; q < MM; ++q) {
        int32_t ys = 0, xs = 0, l = 0, h = 0;
#pragma unroll
        for (int sx = 0; sx < kMfSpecies; ++sx)
          if (sx == spm[k][q]) ys = y[sx], xs = x[sx], l = lo[sx], h = hi[sx];
        iin[r][k][q] = min(max(ys - l, 0), h - l);  // clamped: without a predecessor it adds nothing
        iout[r][k][q] = min(max(xs - l, 0), h - l);
      }
    }
    // the next row's coordinates
    x[0] += 1;
#pragma unroll
    for (int sx = 0; sx + 1 < kMfSpecies; ++sx)
      if (sx + 1 < ns && x[sx] > hi[sx]) {
        x[sx] = lo[sx];
        x[sx + 1] += 1;
      }
  }
  // every load in one batch
  double tin[R][KT][MM], tout[R][KT][MM], gv[R][KT], vi[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    vi[r] = (double)v[min(i032 + r, n32 - 1)];
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      gv[r][k] = (double)v[cc[r][k]];  // in the box: the row itself or its predecessor
#pragma unroll
      for (int q = 0; q < MM; ++q) {
        tin[r][k][q] = tb[k][q][iin[r][k][q]];
        tout[r][k][q] = tb[k][q][iout[r][k][q]];
      }
    }
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (!live[r]) continue;
    double d = 0.0;
    VT wv[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      double ain = 0.0, aout = 0.0;
#pragma unroll
      for (int q = 0; q < MM; ++q) {
        if (q >= cnt[k]) break;
        ain += tin[r][k][q];
        if (leak[r][k])
          d += tout[r][k][q];  // no successor and no sink: the mass leaks
        else if (sink[r][k])
          d += (double)(VT)tout[r][k][q];
        else
          aout += tout[r][k][q];
      }
      if (!leak[r][k] && !sink[r][k]) d += (double)(VT)aout;
      wv[k] = in[r][k] ? (VT)ain : (VT)0;
    }
    double acc = __dmul_rn(-d, vi[r]);
#pragma unroll
    for (int k = 0; k < KT; ++k) acc = __fma_rn((double)wv[k], gv[r][k], acc);
    const int32_t i = i032 + r;
    out[i] = (OT)(accumulate ? (double)out[i] + acc : acc);
  }
}

// The DP5(4) stage combination for rows i .. i + R - 1 (rows >= nw are not written: they
// belong to the sink workgroups), from the stage vector values knew the mat-vec has just
// formed: the same fused multiply-adds in the same order as the separate stage pass, so
// the result is the same to the bit. The k loads are unconditional (a clear mask bit
// points at k_0 again and its term is skipped by a select).
template <typename VT, int R>
__device__ inline void rk_epilogue(const MatArgs<VT>& a, int64_t i, int64_t nw, const double* knew) {
  double kv[6][R], uv[R];
#pragma unroll
  for (int j = 0; j < 6; ++j)
#pragma unroll
    for (int r = 0; r < R; ++r) kv[j][r] = a.rk_k[j][i + r];
#pragma unroll
  for (int r = 0; r < R; ++r) uv[r] = a.rk_u[i + r];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const double t = __fma_rn(a.rk_a[j], kv[j][r], acc);
      acc = ((a.rk_mask >> j) & 1) ? t : acc;
    }
    acc = __fma_rn(a.rk_a[6], knew[r], acc);
    const double y = __fma_rn(a.rk_h, acc, uv[r]);
    if (i + r < nw) {
      a.rk_vin[i + r] = (VT)y;
      if (a.rk_ytmp) a.rk_ytmp[i + r] = y;
    }
  }
}

// Squared scaled DP5(4) error of row i whose new stage value is k7 (ERR kernels): the
// weighted stage sum in stage order, times h, over atol + rtol max(|u|, |y_{n+1}|).
template <typename VT>
__device__ inline double err_row(const MatArgs<VT>& a, int64_t i, double k7) {
  double e = 0.0;
#pragma unroll
  for (int j = 0; j < 6; ++j)
    if (j != 1) e = __fma_rn(a.err_e[j], a.err_k[j][i], e);
  e = __fma_rn(a.err_e[6], k7, e) * a.err_h;
  const double sc = a.err_atol + a.err_rtol * fmax(fabs(a.err_u[i]), fabs(a.err_ynew[i]));
  const double r = e / sc;
  return r * r;
}

// Sink row q of an ERR launch (one thread): its error term and the event's sink values.
template <typename VT>
__device__ inline void err_sink(const MatArgs<VT>& a, int64_t blocks, int q, double k7) {
  const int64_t o = a.n + q;
  a.err_part[blocks + q] = err_row(a, o, k7);
  if (a.err_nex > 0) {
    a.err_ex[q] = a.err_ynew[o];
    a.err_ex[a.nr + q] = a.err_k[0][o];
    a.err_ex[2 * a.nr + q] = k7;
  }
}

// Rows [i0, i0 + R) of the stored generator: out = A v (plus out when accumulating), the
// FP64 row sums left in acc and v at the rows in vi (0 past the states).
template <typename VT, typename OT, int KT, int R, int DM>
__device__ __forceinline__ void stored_rows(const MatArgs<VT>& a, const VT* __restrict__ v, OT* __restrict__ out,
                                            int64_t i0, int accumulate, double (&acc)[R], double (&vi)[R]) {
  // Loads are unconditional (clamped indices, selects) and issued in two batches: the
  // row's own entries and every slot's column and value, then every gather. A load
  // under a branch is waited for at the end of its block, which would make the slots
  // one memory round trip each.
  const Pack<double, R> dg = ldpack<double, R>(a.diag + i0);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const double vr = (double)v[min(i0 + r, a.n - 1)];
    vi[r] = i0 + r < a.n ? vr : 0.0;
  }
  const int K = KT > 0 ? KT : a.K;
  // column of slot k for rows i0..i0+R-1: loaded, or i + dia_off[k] clamped to the
  // states (rows whose value is 0 may point anywhere; they add nothing). DM fixes the
  // kind for every slot (0 loaded, 1 computed) so the fixed-K kernels carry no branch;
  // a matrix with both kinds uses the generic kernel (DM 2), which tests each slot.
  auto slot_cols = [&](int k) {
    Pack<int32_t, R> c;
    const bool dia = DM == 1 || (DM == 2 && ((a.dia_mask >> k) & 1u));
    if (dia) {
      const int64_t hi = a.n - 1;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int64_t j = i0 + r + a.dia_off[k];
        c.v[r] = (int32_t)(j < 0 ? 0 : (j > hi ? hi : j));
      }
    } else {
      c = ldpack<int32_t, R>(a.col + k * a.ld + i0);
    }
    return c;
  };
  if constexpr (KT > 0) {
    // values and columns (computed columns need no load), then every gather; with
    // computed columns the gathers go out with the values, one round trip in all
    Pack<int32_t, R> c[KT];
    Pack<VT, R> w[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      w[k] = ldpack<VT, R>(a.val + k * a.ld + i0);
      c[k] = slot_cols(k);
    }
    double gv[KT][R];
#pragma unroll
    for (int k = 0; k < KT; ++k) {
#pragma unroll
      for (int r = 0; r < R; ++r) gv[k][r] = (double)v[c[k].v[r]];
    }
    __builtin_amdgcn_sched_barrier(0);
    // diag * v_i, then one fused multiply-add per slot in slot order (written out, so
    // that every mode of the kernel rounds alike)
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = __dmul_rn(dg.v[r], vi[r]);
#pragma unroll
    for (int k = 0; k < KT; ++k) {
#pragma unroll
      for (int r = 0; r < R; ++r) acc[r] = __fma_rn((double)w[k].v[r], gv[k][r], acc[r]);
    }
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = __dmul_rn(dg.v[r], vi[r]);
    for (int k = 0; k < K; ++k) {
      const Pack<int32_t, R> c = slot_cols(k);
      const Pack<VT, R> w = ldpack<VT, R>(a.val + k * a.ld + i0);
#pragma unroll
      for (int r = 0; r < R; ++r) acc[r] = __fma_rn((double)w.v[r], (double)v[c.v[r]], acc[r]);
    }
  }
  for (int q = 0; q < a.ntv; ++q) {
    const Pack<VT, R> d = ldpack<VT, R>(a.tvd + q * a.ld + i0);
    const Pack<int32_t, R> c = ldpack<int32_t, R>(a.tvc + q * a.ld + i0);
    const Pack<VT, R> w = ldpack<VT, R>(a.tvv + q * a.ld + i0);
#pragma unroll
    for (int r = 0; r < R; ++r)
      acc[r] += a.tv_scale[q] * ((double)d.v[r] * vi[r] + (double)w.v[r] * (double)v[c.v[r]]);
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t i = i0 + r;
    if (i < a.n) out[i] = (OT)(accumulate ? (double)out[i] + acc[r] : acc[r]);
  }
}

// Thread t owns state rows [R*t, R*t+R). Matrix streams are read with one R-wide vector
// load per slot; v[i] and the predecessor gathers are scalar. EPI: the launch also forms
// the next DP5(4) stage input (rk_epilogue) for its rows, state and sink rows alike.
// DOT: the launch also writes the Arnoldi projection sums of its 512 rows (nc_dual.h):
// the rows' q = A v and v go through LDS to the layout of the block sums, so the
// projection pass need not read q back.
// ERR: the launch also forms the DP5(4) error sums of its rows (err_row), one per row
// workgroup and one per sink row, and the sink values of the event check.
template <typename VT, typename OT, int KT, int R, int DM, bool EPI = false, bool DOT = false, bool ERR = false>
__global__ __launch_bounds__(kBlock) void k_matvec(MatArgs<VT> a, const VT* __restrict__ v, OT* __restrict__ out,
                                                   int accumulate) {
  const int64_t row_threads = (a.n + R - 1) / R;
  const int64_t row_blocks = (row_threads + kBlock - 1) / kBlock;
  if ((int64_t)blockIdx.x < row_blocks) {
    // Workgroups are dealt to the 8 XCDs round-robin; give each XCD one contiguous range
    // of row blocks so the neighbouring v entries a row block gathers sit in its own L2.
    const int64_t b = blockIdx.x, q = row_blocks >> 3, r8 = row_blocks & 7, x = b & 7;
    const int64_t lb = x * q + (x < r8 ? x : r8) + (b >> 3);
    const int64_t t = lb * kBlock + threadIdx.x;
    const int64_t i0 = t * R;
    if constexpr (DOT) {
      static_assert(R * kBlock == kDualRows && DM <= 2 && !EPI, "projection epilogue: 2 rows per thread, stored generator");
      __shared__ double s_x[kDualRows], s_y[kDualRows];
      double vi[R] = {0.0, 0.0}, acc[R] = {0.0, 0.0};
      if (i0 < a.n) stored_rows<VT, OT, KT, R, DM>(a, v, out, i0, accumulate, acc, vi);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const bool live = i0 + r < a.n;  // vi is already 0 past the states
        s_x[R * threadIdx.x + r] = live ? acc[r] : 0.0;
        s_y[R * threadIdx.x + r] = vi[r];
      }
      __syncthreads();
      const int lane = threadIdx.x & 63;
      double xx[8], yy[8];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          xx[4 * h + r] = s_x[256 * h + 4 * lane + r];
          yy[4 * h + r] = s_y[256 * h + 4 * lane + r];
        }
      dual_block<VT>(a.dot_V, a.dot_ld, a.dot_nv, lb * kDualRows, xx, yy, a.dot_sc, a.dot_part, row_blocks, lb);
      return;
    }
    if constexpr (ERR) {
      static_assert(DM <= 2 && !EPI && !DOT, "error epilogue: stored generator");
      double r2 = 0.0;
      if (i0 < a.n) {
        double vi[R], acc[R];
        stored_rows<VT, OT, KT, R, DM>(a, v, out, i0, accumulate, acc, vi);
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (i0 + r < a.n) r2 += err_row(a, i0 + r, acc[r]);
      }
      __shared__ double s_r2[kBlock / 64];
      r2 = wave_sum(r2);
      if ((threadIdx.x & 63) == 0) s_r2[threadIdx.x >> 6] = r2;
      __syncthreads();
      if (threadIdx.x == 0) {
        double sum = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) sum += s_r2[w];
        a.err_part[lb] = sum;
      }
      return;
    }
    if (i0 >= a.n) return;
    if constexpr (DM == 3) {
      mf_rows_values<VT, OT, KT, R>(a, v, out, i0, accumulate);
      return;
    } else if constexpr (DM == 4) {
      mf_rows<VT, OT, KT, R, kMfMembers>(a, v, out, i0, accumulate);
      return;
    }
    double vi[R], acc[R];
    stored_rows<VT, OT, KT, R, DM>(a, v, out, i0, accumulate, acc, vi);
    if constexpr (EPI) rk_epilogue<VT, R>(a, i0, a.n, acc);
  } else {
    const int c = (int)((int64_t)blockIdx.x - row_blocks);
    const int r = a.n_chunks ? a.chunk_row[c] : c;
    const int64_t e0 = a.n_chunks ? a.chunk_start[c] : a.sink_off[r];
    const int64_t e1 = a.n_chunks ? a.chunk_start[c + 1] : a.sink_off[r + 1];
    // four entries per iteration (the thread's entries e, e + 256, ... in the same order as
    // a plain loop), their loads clamped and issued together, then their gathers
    constexpr int U = 4;
    double acc = 0.0;
    for (int64_t eb = e0 + threadIdx.x; eb < e1; eb += (int64_t)kBlock * U) {
      int32_t src[U];
      double w[U], g[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t e = eb + (int64_t)u * kBlock;
        const int64_t ec = e < e1 ? e : e1 - 1;
        src[u] = a.sink_src[ec];
        w[u] = (double)a.sink_val[ec];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) g[u] = (double)v[src[u]];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (eb + (int64_t)u * kBlock < e1) acc += w[u] * g[u];
    }
    __shared__ double part[kBlock / 64];
    __shared__ int s_last;
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) s += part[w];
      if (a.n_chunks) {
        // chunk partial; the chunk workgroup that arrives last (device-scope counter,
        // release fence before the increment, acquire after) folds them below
        __hip_atomic_store(reinterpret_cast<unsigned long long*>(a.chunk_part) + c,
                           (unsigned long long)__double_as_longlong(s), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        const unsigned prev = atomicAdd(a.chunk_arrivals, 1u);
        const int last = prev + 1u == (unsigned)a.n_chunks;
        if (last) __threadfence();
        s_last = last;
      } else {
        s *= a.sink_scale[r];
        const int64_t o = a.n + r;
        out[o] = (OT)(accumulate ? (double)out[o] + s : s);
        if constexpr (EPI) rk_epilogue<VT, 1>(a, o, o + 1, &s);
        if constexpr (ERR) err_sink(a, row_blocks, r, s);
      }
    }
    if (a.n_chunks) {
      __syncthreads();
      if (s_last) {
        // sink row q = sum of its chunk partials (lane l takes chunks l, l + 64, ..., then a
        // fixed shuffle tree: deterministic), times its time factor; one wave per row
        const int lane = threadIdx.x & 63;
        for (int q = threadIdx.x >> 6; q < a.nr; q += kBlock / 64) {
          double s = 0.0;
          for (int k = a.chunk_first[q] + lane; k < a.chunk_first[q + 1]; k += 64)
            s += __longlong_as_double((long long)__hip_atomic_load(reinterpret_cast<unsigned long long*>(a.chunk_part) + k,
                                                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
          s = wave_sum(s);
          if (lane == 0) {
            s *= a.sink_scale[q];
            const int64_t o = a.n + q;
            out[o] = (OT)(accumulate ? (double)out[o] + s : s);
            if constexpr (EPI) rk_epilogue<VT, 1>(a, o, o + 1, &s);
            if constexpr (ERR) err_sink(a, row_blocks, q, s);
          }
        }
        if (threadIdx.x == 0) __hip_atomic_store(a.chunk_arrivals, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
}

// Launch timing: when e0/e1 are given the kernel is launched with hipExtLaunchKernelGGL,
// which stamps the events from the dispatch itself (the kernel's own start and end), so
// the recorded durations carry no event-packet or dispatch gaps.
struct LaunchCfg {
  unsigned grid;
  hipStream_t s;
  hipEvent_t e0, e1;
};

template <typename VT, typename OT, int K, int R, int DM, bool EPI = false, bool DOT = false, bool ERR = false>
void launch_one(const LaunchCfg& c, const MatArgs<VT>& a, const VT* vv, OT* oo, int acc) {
  if (c.e0)
    hipExtLaunchKernelGGL(k_matvec<VT, OT, K, R, DM, EPI, DOT, ERR>, dim3(c.grid), dim3(kBlock), 0, c.s, c.e0, c.e1, 0, a,
                          vv, oo, acc);
  else
    k_matvec<VT, OT, K, R, DM, EPI, DOT, ERR><<<c.grid, kBlock, 0, c.s>>>(a, vv, oo, acc);
}

// The error-epilogue kernels: FP64 output, 2 rows per thread, stored generator
template <typename VT, int DM>
void dispatch_err_kd(int K, const LaunchCfg& c, const MatArgs<VT>& a, const VT* vv, double* oo) {
  switch (K) {
    case 1: launch_one<VT, double, 1, 2, DM, false, false, true>(c, a, vv, oo, 0); break;
    case 2: launch_one<VT, double, 2, 2, DM, false, false, true>(c, a, vv, oo, 0); break;
    case 3: launch_one<VT, double, 3, 2, DM, false, false, true>(c, a, vv, oo, 0); break;
    case 4: launch_one<VT, double, 4, 2, DM, false, false, true>(c, a, vv, oo, 0); break;
    case 5: launch_one<VT, double, 5, 2, DM, false, false, true>(c, a, vv, oo, 0); break;
    case 6: launch_one<VT, double, 6, 2, DM, false, false, true>(c, a, vv, oo, 0); break;
    case 7: launch_one<VT, double, 7, 2, DM, false, false, true>(c, a, vv, oo, 0); break;
    case 8: launch_one<VT, double, 8, 2, DM, false, false, true>(c, a, vv, oo, 0); break;
    default: launch_one<VT, double, 0, 2, 2, false, false, true>(c, a, vv, oo, 0); break;
  }
}

// The projection-epilogue kernels: FP64 output, 2 rows per thread, every slot's column
// loaded (DM 0) or every slot's computed (DM 1), at most 8 slots
template <typename VT, int DM>
void dispatch_dots_kd(int K, const LaunchCfg& c, const MatArgs<VT>& a, const VT* vv, double* oo) {
  switch (K) {
    case 1: launch_one<VT, double, 1, 2, DM, false, true>(c, a, vv, oo, 0); break;
    case 2: launch_one<VT, double, 2, 2, DM, false, true>(c, a, vv, oo, 0); break;
    case 3: launch_one<VT, double, 3, 2, DM, false, true>(c, a, vv, oo, 0); break;
    case 4: launch_one<VT, double, 4, 2, DM, false, true>(c, a, vv, oo, 0); break;
    case 5: launch_one<VT, double, 5, 2, DM, false, true>(c, a, vv, oo, 0); break;
    case 6: launch_one<VT, double, 6, 2, DM, false, true>(c, a, vv, oo, 0); break;
    case 7: launch_one<VT, double, 7, 2, DM, false, true>(c, a, vv, oo, 0); break;
    default: launch_one<VT, double, 8, 2, DM, false, true>(c, a, vv, oo, 0); break;
  }
}

// The epilogue kernels: FP64 output, 2 rows per thread, stored generator (loaded, computed
// or mixed columns)
template <typename VT, int DM>
void dispatch_epi_kd(int K, const LaunchCfg& c, const MatArgs<VT>& a, const VT* vv, double* oo) {
  switch (K) {
    case 1: launch_one<VT, double, 1, 2, DM, true>(c, a, vv, oo, 0); break;
    case 2: launch_one<VT, double, 2, 2, DM, true>(c, a, vv, oo, 0); break;
    case 3: launch_one<VT, double, 3, 2, DM, true>(c, a, vv, oo, 0); break;
    case 4: launch_one<VT, double, 4, 2, DM, true>(c, a, vv, oo, 0); break;
    case 5: launch_one<VT, double, 5, 2, DM, true>(c, a, vv, oo, 0); break;
    case 6: launch_one<VT, double, 6, 2, DM, true>(c, a, vv, oo, 0); break;
    case 7: launch_one<VT, double, 7, 2, DM, true>(c, a, vv, oo, 0); break;
    case 8: launch_one<VT, double, 8, 2, DM, true>(c, a, vv, oo, 0); break;
    default: launch_one<VT, double, 0, 2, 2, true>(c, a, vv, oo, 0); break;
  }
}

template <typename VT>
void dispatch_epi(int K, const LaunchCfg& c, const MatArgs<VT>& a, const VT* vv, double* oo) {
  const uint32_t all = K >= 32 ? 0xffffffffu : ((1u << K) - 1u);
  if (a.dia_mask == 0u)
    dispatch_epi_kd<VT, 0>(K, c, a, vv, oo);
  else if ((a.dia_mask & all) == all)
    dispatch_epi_kd<VT, 1>(K, c, a, vv, oo);
  else
    launch_one<VT, double, 0, 2, 2, true>(c, a, vv, oo, 0);
}

template <typename VT, typename OT, int R, int DM>
void dispatch_kd(int K, const LaunchCfg& c, const MatArgs<VT>& a, const VT* vv, OT* oo, int acc) {
  switch (K) {
    case 1: launch_one<VT, OT, 1, R, DM>(c, a, vv, oo, acc); break;
    case 2: launch_one<VT, OT, 2, R, DM>(c, a, vv, oo, acc); break;
    case 3: launch_one<VT, OT, 3, R, DM>(c, a, vv, oo, acc); break;
    case 4: launch_one<VT, OT, 4, R, DM>(c, a, vv, oo, acc); break;
    case 5: launch_one<VT, OT, 5, R, DM>(c, a, vv, oo, acc); break;
    case 6: launch_one<VT, OT, 6, R, DM>(c, a, vv, oo, acc); break;
    case 7: launch_one<VT, OT, 7, R, DM>(c, a, vv, oo, acc); break;
    case 8: launch_one<VT, OT, 8, R, DM>(c, a, vv, oo, acc); break;
    default: launch_one<VT, OT, 0, R, 2>(c, a, vv, oo, acc); break;
  }
}

// Column mode: every slot loaded (0) or computed (1) takes the fixed-K kernels; a mix of
// the two takes the generic kernel, which decides per slot.
template <typename VT, typename OT, int R>
void dispatch_k(int K, const LaunchCfg& c, const MatArgs<VT>& a, const VT* vv, OT* oo, int acc) {
  const uint32_t all = K >= 32 ? 0xffffffffu : ((1u << K) - 1u);
  if (a.dia_mask == 0u)
    dispatch_kd<VT, OT, R, 0>(K, c, a, vv, oo, acc);
  else if ((a.dia_mask & all) == all)
    dispatch_kd<VT, OT, R, 1>(K, c, a, vv, oo, acc);
  else
    launch_one<VT, OT, 0, R, 2>(c, a, vv, oo, acc);
}

// Launches of the computed modes (FP64 output, 1 or 2 rows per thread; DM 3: computed
// values and the stored diagonal; DM 4: fully matrix-free).
template <typename VT, int R, int DM>
void dispatch_mf(int K, const LaunchCfg& c, const MatArgs<VT>& a, const VT* vv, double* oo, int acc) {
  switch (K) {
    case 1: launch_one<VT, double, 1, R, DM>(c, a, vv, oo, acc); break;
    case 2: launch_one<VT, double, 2, R, DM>(c, a, vv, oo, acc); break;
    case 3: launch_one<VT, double, 3, R, DM>(c, a, vv, oo, acc); break;
    case 4: launch_one<VT, double, 4, R, DM>(c, a, vv, oo, acc); break;
    case 5: launch_one<VT, double, 5, R, DM>(c, a, vv, oo, acc); break;
    case 6: launch_one<VT, double, 6, R, DM>(c, a, vv, oo, acc); break;
    case 7: launch_one<VT, double, 7, R, DM>(c, a, vv, oo, acc); break;
    default: launch_one<VT, double, 8, R, DM>(c, a, vv, oo, acc); break;
  }
}

template <typename VT>
void dispatch_mf_r(int R, bool full, int K, const LaunchCfg& c, const MatArgs<VT>& a, const VT* vv, double* oo,
                   int acc) {
  // full: the matrix-free form (DM 4); otherwise computed values with the stored diagonal (DM 3)
  if (R == 1)
    full ? dispatch_mf<VT, 1, 4>(K, c, a, vv, oo, acc) : dispatch_mf<VT, 1, 3>(K, c, a, vv, oo, acc);
  else
    full ? dispatch_mf<VT, 2, 4>(K, c, a, vv, oo, acc) : dispatch_mf<VT, 2, 3>(K, c, a, vv, oo, acc);
}

template <typename VT, typename OT>
void dispatch_r(int R, int K, const LaunchCfg& c, const MatArgs<VT>& a, const VT* vv, OT* oo, int acc) {
  if (R == 4)
    dispatch_k<VT, OT, 4>(K, c, a, vv, oo, acc);
  else if (R == 2)
    dispatch_k<VT, OT, 2>(K, c, a, vv, oo, acc);
  else
    dispatch_k<VT, OT, 1>(K, c, a, vv, oo, acc);
}

template <typename VT>
void launch_matvec_t(nc_matrix_s* A, double t, const void* v, void* out, bool accumulate, hipStream_t s,
                     bool out_fp64, hipEvent_t e0, hipEvent_t e1, const RkEpilogue* epi = nullptr,
                     const DualDots* dots = nullptr, const RkErrEpilogue* err = nullptr) {
  MatArgs<VT> a{};
  a.n = A->n;
  a.ld = A->ld;
  a.K = A->K;
  a.ntv = (int32_t)A->tv_r.size();
  a.nr = A->nr;
  a.diag = A->diag.p;
  a.col = A->col.p;
  a.val = reinterpret_cast<const VT*>(A->val.p);
  a.tvd = reinterpret_cast<const VT*>(A->tvd.p);
  a.tvc = A->tvc.p;
  a.tvv = reinterpret_cast<const VT*>(A->tvv.p);
  a.sink_off = A->sink_off.p;
  a.sink_src = A->sink_src.p;
  a.sink_val = reinterpret_cast<const VT*>(A->sink_val.p);
  a.dia_mask = A->dia_mask;
  a.n_chunks = A->n_chunks;
  a.chunk_start = A->chunk_start.p;
  a.chunk_row = A->chunk_row.p;
  a.chunk_part = A->chunk_part.p;
  a.chunk_first = A->chunk_first.p;
  a.chunk_arrivals = A->chunk_arrivals.p;
  for (size_t k = 0; k < A->dia_off.size() && k < (size_t)kMaxReactions; ++k) a.dia_off[k] = A->dia_off[k];
  a.mf = A->mf_mode ? reinterpret_cast<const MfInfo*>(A->mf_info.p) : nullptr;
  a.mf_tab = A->mf_mode ? A->mf_tab.p : nullptr;
  const nc_model_s* m = A->model;
  for (int r = 0; r < A->nr; ++r) a.sink_scale[r] = m->kinds[r] == NC_PROP_SEPARABLE ? m->tfactor(r, t) : 1.0;
  for (size_t q = 0; q < A->tv_r.size(); ++q)
    a.tv_scale[q] = A->tv_kind[q] == NC_PROP_SEPARABLE ? m->tfactor(A->tv_r[q], t) : 1.0;
  // the matrix-free kernel takes 1 or 2 rows per thread (4 runs as 2)
  const int R = (out_fp64 && A->mf_mode && A->rows_per_thread == 4) ? 2 : A->rows_per_thread;
  const int64_t row_threads = (A->n + R - 1) / R;
  const int64_t row_blocks = (row_threads + kBlock - 1) / kBlock;
  const LaunchCfg c{(unsigned)(row_blocks + (A->n_chunks ? A->n_chunks : A->nr)), s, e0, e1};
  const VT* vv = reinterpret_cast<const VT*>(v);
  const int acc = accumulate ? 1 : 0;
  if (epi) {
    if (!matvec_rk_fusable(A) || accumulate || !out_fp64) throw Error(NC_ERR_INTERNAL, "mat-vec epilogue not available");
    for (int j = 0; j < 6; ++j) {
      const bool on = (epi->mask >> j) & 1;
      a.rk_k[j] = on ? epi->k[j] : epi->k[0];
      a.rk_a[j] = on ? epi->a[j] : 0.0;
    }
    a.rk_a[6] = epi->a_new;
    a.rk_h = epi->h;
    a.rk_u = epi->u;
    a.rk_ytmp = epi->ytmp;
    a.rk_vin = reinterpret_cast<VT*>(epi->vin);
    a.rk_mask = epi->mask;
    dispatch_epi<VT>(A->K, c, a, vv, reinterpret_cast<double*>(out));
    return;
  }
  if (err) {
    if (!matvec_rk_fusable(A) || accumulate || !out_fp64 || R != 2)
      throw Error(NC_ERR_INTERNAL, "mat-vec error epilogue not available");
    for (int j = 0; j < 6; ++j) a.err_k[j] = err->k[j] ? err->k[j] : err->k[0];
    for (int j = 0; j < 7; ++j) a.err_e[j] = err->e[j];
    a.err_h = err->h;
    a.err_atol = err->atol;
    a.err_rtol = err->rtol;
    a.err_u = err->u;
    a.err_ynew = err->ynew;
    a.err_part = err->part;
    a.err_ex = err->ex;
    a.err_nex = err->n_ex;
    const uint32_t all = A->K >= 32 ? 0xffffffffu : ((1u << A->K) - 1u);
    if (a.dia_mask == 0u)
      dispatch_err_kd<VT, 0>(A->K, c, a, vv, reinterpret_cast<double*>(out));
    else if ((a.dia_mask & all) == all)
      dispatch_err_kd<VT, 1>(A->K, c, a, vv, reinterpret_cast<double*>(out));
    else
      launch_one<VT, double, 0, 2, 2, false, false, true>(c, a, vv, reinterpret_cast<double*>(out), 0);
    return;
  }
  if (dots) {
    if (!matvec_dots_fusable(A) || accumulate || !out_fp64 || R != 2)
      throw Error(NC_ERR_INTERNAL, "mat-vec projection epilogue not available");
    if (dots->nv < 1 || dots->nv > 32 || dots->nblk != row_blocks)
      throw Error(NC_ERR_INTERNAL, "projection blocks do not match the mat-vec's row workgroups");
    a.dot_V = reinterpret_cast<const VT*>(dots->V);
    a.dot_ld = dots->ld;
    a.dot_nv = dots->nv;
    a.dot_sc = dots->sc;
    a.dot_part = dots->part;
    if (a.dia_mask == 0u)
      dispatch_dots_kd<VT, 0>(A->K, c, a, vv, reinterpret_cast<double*>(out));
    else
      dispatch_dots_kd<VT, 1>(A->K, c, a, vv, reinterpret_cast<double*>(out));
    return;
  }
  if (out_fp64 && A->mf_mode)
    dispatch_mf_r<VT>(R, A->mf_mode == 2, A->K, c, a, vv, reinterpret_cast<double*>(out), acc);
  else if (out_fp64)
    dispatch_r<VT, double>(R, A->K, c, a, vv, reinterpret_cast<double*>(out), acc);
  else
    dispatch_r<VT, VT>(R, A->K, c, a, vv, reinterpret_cast<VT*>(out), acc);
}

// ---- matrix-free setup --------------------------------------------------------------
// lohi[2s] = min, lohi[2s + 1] = max of species s over the states (grid-stride; one atomic
// pair per species and workgroup)
__global__ __launch_bounds__(kBlock) void k_mf_bounds(int64_t n, int ns, const int32_t* states, int32_t* lohi) {
  __shared__ int32_t slo[kMfSpecies][kBlock / 64], shi[kMfSpecies][kBlock / 64];
  int32_t lo[kMfSpecies], hi[kMfSpecies];
  for (int sx = 0; sx < kMfSpecies; ++sx) lo[sx] = INT32_MAX, hi[sx] = INT32_MIN;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    for (int sx = 0; sx < kMfSpecies; ++sx)
      if (sx < ns) {
        const int32_t v = states[i * ns + sx];
        lo[sx] = min(lo[sx], v);
        hi[sx] = max(hi[sx], v);
      }
  for (int sx = 0; sx < kMfSpecies; ++sx) {
    for (int off = 32; off > 0; off >>= 1) {
      lo[sx] = min(lo[sx], __shfl_xor(lo[sx], off, 64));
      hi[sx] = max(hi[sx], __shfl_xor(hi[sx], off, 64));
    }
    if ((threadIdx.x & 63) == 0) slo[sx][threadIdx.x >> 6] = lo[sx], shi[sx][threadIdx.x >> 6] = hi[sx];
  }
  __syncthreads();
  if (threadIdx.x < (unsigned)ns) {
    const int sx = threadIdx.x;
    int32_t l = slo[sx][0], h = shi[sx][0];
    for (int w = 1; w < kBlock / 64; ++w) l = min(l, slo[sx][w]), h = max(h, shi[sx][w]);
    atomicMin(lohi + 2 * sx, l);
    atomicMax(lohi + 2 * sx + 1, h);
  }
}

// *bad = 1 when some state is not at its lattice position (species 0 fastest)
__global__ void k_mf_check(int64_t n, const int32_t* states, const MfInfo* mi, unsigned* bad) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int ns = mi->ns;
  uint64_t rem = (uint64_t)i;
  bool ok = true;
  for (int sx = 0; sx < ns; ++sx) {
    const uint64_t e = (uint64_t)(mi->hi[sx] - mi->lo[sx] + 1);
    ok = ok && states[i * ns + sx] == mi->lo[sx] + (int32_t)(rem % e);
    rem /= e;
  }
  if (!ok) atomicOr(bad, 1u);
}

// Entry e of the tables: reaction r's propensity at coordinate lo + (e - toff[r]) of its
// species (every other species at its lower bound; the program does not read them), or its
// constant value (one entry) when it reads no species.
__global__ void k_mf_tables(int64_t total, int nr, const MfInfo* mi, Progs pr, double* tab) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= total) return;
  int r = 0;
  while (r + 1 < nr && mi->toff[r + 1] <= e) ++r;
  int32_t x[kMaxSpecies];
  for (int sx = 0; sx < mi->ns; ++sx) x[sx] = mi->lo[sx];
  const int spm = mi->sp[r];
  if (spm >= 0) x[spm] = mi->lo[spm] + (int32_t)(e - mi->toff[r]);
  tab[e] = xprog(pr, r, 0.0, x);
}

// Decides matrix-free eligibility after the stored build and, when eligible, builds the
// tables and turns the mode on.
void mf_setup(nc_matrix_s* A, nc_space_s* sp, nc_model_s* m, const Groups& g) {
  A->mf_ok = false;
  A->mf_mode = 0;
  const int ns = sp->ns, nr = sp->nr;
  const int64_t n = A->n;
  if (n == 0 || n >= (int64_t)INT32_MAX || ns > kMfSpecies || A->K != g.ng || A->K < 1 || A->K > kMfSlots ||
      !A->tv_r.empty())
    return;
  // a box in lattice order has every slot's column at one offset (checked by the build's
  // column detection): anything else needs no further look
  const uint32_t all_slots = A->K >= 32 ? 0xffffffffu : ((1u << A->K) - 1u);
  if ((A->dia_mask & all_slots) != all_slots) return;
  A->mf_multi = false;
  for (int k = 0; k < g.ng; ++k) {
    if (g.start[k + 1] - g.start[k] > kMfMembers) return;
    A->mf_multi = A->mf_multi || g.start[k + 1] - g.start[k] > 1;
  }
  // every propensity reads at most one species
  std::vector<int32_t> spr(nr, -1);
  for (int r = 0; r < nr; ++r) {
    const nc_instr* c = m->xprog(r);
    for (int64_t k = 0; k < m->xlen(r); ++k) {
      if (c[k].op != NC_OP_SPECIES) continue;
      if (spr[r] < 0)
        spr[r] = c[k].arg;
      else if (spr[r] != c[k].arg)
        return;
    }
  }
  hipStream_t s = A->stream.s;
  DevBuf<int32_t> lohi(2 * kMfSpecies);
  int32_t* h = static_cast<int32_t*>(host_stage(5));
  for (int sx = 0; sx < kMfSpecies; ++sx) h[2 * sx] = INT32_MAX, h[2 * sx + 1] = INT32_MIN;
  NC_CHECK_HIP(hipMemcpyAsync(lohi.p, h, sizeof(int32_t) * 2 * kMfSpecies, hipMemcpyHostToDevice, s));
  k_mf_bounds<<<std::min<unsigned>(grid_for(n, kBlock), 1024), kBlock, 0, s>>>(n, ns, sp->states.p, lohi.p);
  NC_CHECK_HIP(hipGetLastError());
  NC_CHECK_HIP(hipMemcpyAsync(h, lohi.p, sizeof(int32_t) * 2 * kMfSpecies, hipMemcpyDeviceToHost, s));
  NC_CHECK_HIP(hipStreamSynchronize(s));
  MfInfo info{};
  info.ns = ns;
  info.ng = g.ng;
  double cells = 1.0;
  int64_t stride = 1;
  for (int sx = 0; sx < ns; ++sx) {
    info.lo[sx] = h[2 * sx];
    info.hi[sx] = h[2 * sx + 1];
    info.stride[sx] = stride;
    const int64_t e = (int64_t)info.hi[sx] - info.lo[sx] + 1;
    cells *= (double)e;
    stride *= e;
  }
  if (cells != (double)n) return;  // not a full box
  for (int k = 0; k <= g.ng; ++k) info.start[k] = g.start[k];
  for (int q = 0; q < g.start[g.ng]; ++q) info.members[q] = g.members[q];
  for (int k = 0; k < g.ng; ++k) {
    int64_t off = 0;
    int changed = 0;
    info.cs[k] = 0;
    for (int sx = 0; sx < ns; ++sx) {
      info.nu[k][sx] = m->stoich[sx + g.leader[k] * ns];
      off += (int64_t)info.nu[k][sx] * info.stride[sx];
      if (info.nu[k][sx] != 0) {
        info.cs[k] = sx;
        ++changed;
      }
    }
    info.off[k] = off;
    // the lean form needs one changed species per slot (and one reaction per slot)
    if (changed != 1) A->mf_multi = true;
  }
  int64_t total = 0;
  for (int r = 0; r < nr; ++r) {
    info.sp[r] = spr[r];
    info.toff[r] = total;
    total += spr[r] < 0 ? 1 : (int64_t)info.hi[spr[r]] - info.lo[spr[r]] + 1;
  }
  A->mf_info.alloc(sizeof(MfInfo));
  NC_CHECK_HIP(hipMemcpyAsync(A->mf_info.p, &info, sizeof(MfInfo), hipMemcpyHostToDevice, s));
  DevBuf<unsigned> bad(1);
  NC_CHECK_HIP(hipMemsetAsync(bad.p, 0, sizeof(unsigned), s));
  const MfInfo* dinfo = reinterpret_cast<const MfInfo*>(A->mf_info.p);
  k_mf_check<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, sp->states.p, dinfo, bad.p);
  A->mf_tab.alloc((size_t)total);
  const Progs pr = make_progs(m);
  k_mf_tables<<<grid_for(total, kBlock), kBlock, 0, s>>>(total, nr, dinfo, pr, A->mf_tab.p);
  NC_CHECK_HIP(hipGetLastError());
  unsigned* hb = reinterpret_cast<unsigned*>(h);
  NC_CHECK_HIP(hipMemcpyAsync(hb, bad.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
  NC_CHECK_HIP(hipStreamSynchronize(s));  // info (a stack copy) and the tables are complete
  if (*hb) {
    A->mf_tab.release();
    A->mf_info.release();
    return;  // not in lattice order
  }
  A->mf_table_entries = total;
  A->mf_ok = true;
  // opt-in: on MI355X the stored kernel (columns computed, values and diagonal streamed) is
  // faster than both computed forms, which are bound by their index and select arithmetic
  // (C2's 1024^2 toggle: 8.9 us stored, 12.7 us computed values, 41 us matrix-free)
  A->mf_mode = 0;
}

template <typename VT>
void joint_refresh(nc_matrix_s* A, double t) {
  bool any = false;
  for (int k : A->tv_kind) any = any || k == NC_PROP_JOINT;
  if (!any || t == A->t_cache) return;
  A->t_cache = t;
  const nc_model_s* m = A->model;
  TvList tv{};
  tv.n = (int32_t)A->tv_r.size();
  for (int q = 0; q < tv.n; ++q) tv.r[q] = A->tv_r[q], tv.kind[q] = A->tv_kind[q];
  const Progs pr = make_progs(m);
  hipStream_t s = A->stream.s;
  if (A->n)
    k_joint_refresh<VT><<<grid_for(A->n, kBlock), kBlock, 0, s>>>(
        A->n, A->ld, A->ns, A->nr, A->states.p, A->state_conn.p, m->d_stoich.p, tv, pr, t,
        reinterpret_cast<VT*>(A->tvd.p), reinterpret_cast<VT*>(A->tvv.p));
  for (int r = 0; r < A->nr; ++r) {
    if (m->kinds[r] != NC_PROP_JOINT) continue;
    const int64_t e0 = A->sink_off_h[r], e1 = A->sink_off_h[r + 1];
    if (e1 > e0)
      k_joint_sink_refresh<VT><<<grid_for(e1 - e0, kBlock), kBlock, 0, s>>>(e0, e1, r, A->ns, A->states.p,
                                                                             A->sink_src.p, pr, t,
                                                                             reinterpret_cast<VT*>(A->sink_val.p));
  }
  NC_CHECK_HIP(hipGetLastError());
}

template <typename VT>
void build_t(nc_matrix_s* A, nc_space_s* sp, nc_model_s* m) {
  hipStream_t s = A->stream.s;
  const int64_t n = sp->n;
  const int nr = sp->nr, ns = sp->ns;
  const int64_t ld = ((n + kLdAlign - 1) / kLdAlign) * kLdAlign;
  A->ld = ld;
  // groups of time-invariant reactions with identical stoichiometry (merged by sparse())
  Groups g{};
  std::vector<std::vector<int>> groups;
  for (int r = 0; r < nr; ++r) {
    if (m->kinds[r] != NC_PROP_TIME_INVARIANT) continue;
    bool placed = false;
    for (auto& gr : groups) {
      const int l = gr[0];
      bool same = true;
      for (int sidx = 0; sidx < ns; ++sidx) same = same && m->stoich[sidx + l * ns] == m->stoich[sidx + r * ns];
      if (same) {
        gr.push_back(r);
        placed = true;
        break;
      }
    }
    if (!placed) groups.push_back({r});
  }
  g.ng = (int32_t)groups.size();
  int mcount = 0;
  for (int k = 0; k < g.ng; ++k) {
    g.leader[k] = groups[k][0];
    g.start[k] = mcount;
    for (int r : groups[k]) g.members[mcount++] = r;
  }
  g.start[g.ng] = mcount;
  TvList tv{};
  for (int r = 0; r < nr; ++r)
    if (m->kinds[r] != NC_PROP_TIME_INVARIANT) {
      tv.r[tv.n] = r;
      tv.kind[tv.n] = m->kinds[r];
      A->tv_r.push_back(r);
      A->tv_kind.push_back(m->kinds[r]);
      ++tv.n;
    }
  // pass 1: ELL width and counts
  DevBuf<int32_t> d_kmax(1);
  DevBuf<unsigned long long> d_counts(2 + kMaxReactions);
  NC_CHECK_HIP(hipMemsetAsync(d_kmax.p, 0, sizeof(int32_t), s));
  NC_CHECK_HIP(hipMemsetAsync(d_counts.p, 0, sizeof(unsigned long long) * (2 + kMaxReactions), s));
  if (n)
    k_build_count<<<std::min<unsigned>(grid_for(n, kBlock), 1024), kBlock, 0, s>>>(n, nr, sp->state_conn.p, g, tv, d_kmax.p, d_counts.p,
                                                         d_counts.p + 2);
  NC_CHECK_HIP(hipGetLastError());
  unsigned long long counts[2 + kMaxReactions];
  {
    unsigned char* h = static_cast<unsigned char*>(host_stage(1));
    NC_CHECK_HIP(hipMemcpyAsync(h, d_counts.p, sizeof(counts), hipMemcpyDeviceToHost, s));
    NC_CHECK_HIP(hipMemcpyAsync(h + sizeof(counts), d_kmax.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    NC_CHECK_HIP(hipStreamSynchronize(s));
    std::memcpy(counts, h, sizeof(counts));
    std::memcpy(&A->K, h + sizeof(counts), sizeof(int32_t));
  }
  const int32_t K = A->K;
  const size_t vs = sizeof(VT);
  A->diag.alloc(std::max<int64_t>(ld, 1));
  A->col.alloc(std::max<int64_t>((int64_t)K * ld, 1));
  A->val.alloc(std::max<int64_t>((int64_t)K * ld, 1) * vs);
  A->tvd.alloc(std::max<int64_t>((int64_t)tv.n * ld, 1) * vs);
  A->tvv.alloc(std::max<int64_t>((int64_t)tv.n * ld, 1) * vs);
  A->tvc.alloc(std::max<int64_t>((int64_t)tv.n * ld, 1));
  const Progs pr = make_progs(m);
  DevBuf<unsigned long long> d_dmax(1);
  NC_CHECK_HIP(hipMemsetAsync(d_dmax.p, 0, sizeof(unsigned long long), s));
  if (ld) {
    auto fill = pr.reg ? k_build_fill<VT, true> : k_build_fill<VT, false>;
    fill<<<grid_for(ld, kBlock), kBlock, 0, s>>>(
        n, ld, ns, nr, K, K == g.ng ? 1 : 0, sp->states.p, sp->state_conn.p, sp->sink_mask.p, sp->d_stoich.p, g, tv, pr,
        A->diag.p, A->col.p, reinterpret_cast<VT*>(A->val.p), reinterpret_cast<VT*>(A->tvd.p), A->tvc.p,
        reinterpret_cast<VT*>(A->tvv.p), d_dmax.p);
  }
  NC_CHECK_HIP(hipGetLastError());
  // slots with one column offset in every nonzero row (read back with the final sync)
  const int kd = std::min<int>(K, kMaxReactions);
  // pinned: the DIA results and dmax land here with the final synchronisation
  int32_t* dia_h = static_cast<int32_t*>(host_stage(3));
  unsigned long long* dmax_h = reinterpret_cast<unsigned long long*>(dia_h + 2 * kMaxReactions);
  DevBuf<int32_t> d_dia(2 * std::max(kd, 1));
  if (n && kd) {
    std::vector<int32_t> dia_init(2 * kd);
    for (int k = 0; k < kd; ++k) dia_init[2 * k] = INT32_MAX, dia_init[2 * k + 1] = INT32_MIN;
    NC_CHECK_HIP(hipMemcpyAsync(d_dia.p, dia_init.data(), sizeof(int32_t) * 2 * kd, hipMemcpyHostToDevice, s));
    k_dia_detect<VT><<<dim3(std::min<unsigned>(grid_for(n, kBlock), 256), (unsigned)kd), kBlock, 0, s>>>(
        n, ld, A->col.p, reinterpret_cast<const VT*>(A->val.p), d_dia.p);
    NC_CHECK_HIP(hipGetLastError());
    NC_CHECK_HIP(hipMemcpyAsync(dia_h, d_dia.p, sizeof(int32_t) * 2 * kd, hipMemcpyDeviceToHost, s));
  }
  // sink entries, grouped by sink row (reaction), ascending source state within a row
  DevBuf<uint32_t> flags(std::max<int64_t>(n * nr, 1)), pos(std::max<int64_t>(n * nr, 1));
  DevBuf<unsigned char> scan_tmp;
  uint32_t total = 0;
  if (n) {
    k_sink_flags<<<grid_for(n * nr, kBlock), kBlock, 0, s>>>(n, nr, sp->sink_mask.p, flags.p);
    exclusive_scan_u32(flags.p, pos.p, n * nr, scan_tmp, s, &total);
  }
  A->sink_off_h.assign(nr + 1, 0);
  if (n) {
    // the first entry of each sink row (pos at r * n), gathered on the device and read back
    // in one copy
    DevBuf<uint32_t> d_starts(nr);
    k_gather_strided<<<1, 64, 0, s>>>(pos.p, n, nr, d_starts.p);
    NC_CHECK_HIP(hipGetLastError());
    uint32_t* starts = static_cast<uint32_t*>(host_stage(2));
    NC_CHECK_HIP(hipMemcpyAsync(starts, d_starts.p, sizeof(uint32_t) * nr, hipMemcpyDeviceToHost, s));
    NC_CHECK_HIP(hipStreamSynchronize(s));
    for (int r = 0; r < nr; ++r) A->sink_off_h[r] = starts[r];
  }
  A->sink_off_h[nr] = total;
  A->n_sink_entries = total;
  // long sink rows (above 8192 entries): 4096-entry chunks, each reduced by its own
  // workgroup (the host vectors stay alive until the build's final synchronisation)
  constexpr int64_t kSinkChunk = 4096, kChunkFrom = 8192;
  std::vector<int64_t> chunk_start_h;
  std::vector<int32_t> chunk_row_h, chunk_first_h;
  {
    int64_t longest = 0;
    for (int r = 0; r < nr; ++r) longest = std::max<int64_t>(longest, A->sink_off_h[r + 1] - A->sink_off_h[r]);
    A->n_chunks = 0;
    if (longest > kChunkFrom) {
      chunk_first_h.assign(nr + 1, 0);
      for (int r = 0; r < nr; ++r) {
        chunk_first_h[r] = (int32_t)chunk_row_h.size();
        for (int64_t e = A->sink_off_h[r]; e < A->sink_off_h[r + 1]; e += kSinkChunk) {
          chunk_start_h.push_back(e);
          chunk_row_h.push_back(r);
        }
      }
      chunk_first_h[nr] = (int32_t)chunk_row_h.size();
      chunk_start_h.push_back(total);
      A->n_chunks = (int32_t)chunk_row_h.size();
      A->chunk_start.alloc(chunk_start_h.size());
      A->chunk_row.alloc(chunk_row_h.size());
      A->chunk_first.alloc(chunk_first_h.size());
      A->chunk_part.alloc(chunk_row_h.size());
      A->chunk_arrivals.alloc(1);
      NC_CHECK_HIP(hipMemsetAsync(A->chunk_arrivals.p, 0, sizeof(unsigned), s));
      NC_CHECK_HIP(hipMemcpyAsync(A->chunk_start.p, chunk_start_h.data(), sizeof(int64_t) * chunk_start_h.size(),
                                  hipMemcpyHostToDevice, s));
      NC_CHECK_HIP(hipMemcpyAsync(A->chunk_row.p, chunk_row_h.data(), sizeof(int32_t) * chunk_row_h.size(),
                                  hipMemcpyHostToDevice, s));
      NC_CHECK_HIP(hipMemcpyAsync(A->chunk_first.p, chunk_first_h.data(), sizeof(int32_t) * chunk_first_h.size(),
                                  hipMemcpyHostToDevice, s));
    }
  }
  A->sink_off.alloc(nr + 1);
  NC_CHECK_HIP(hipMemcpyAsync(A->sink_off.p, A->sink_off_h.data(), sizeof(int64_t) * (nr + 1), hipMemcpyHostToDevice, s));
  A->sink_src.alloc(std::max<uint32_t>(total, 1));
  A->sink_val.alloc(std::max<uint32_t>(total, 1) * vs);
  DevBuf<int32_t> d_kinds(nr);
  NC_CHECK_HIP(hipMemcpyAsync(d_kinds.p, m->kinds.data(), sizeof(int32_t) * nr, hipMemcpyHostToDevice, s));
  if (n && total) {
    auto sfill = pr.reg ? k_sink_fill<VT, true> : k_sink_fill<VT, false>;
    sfill<<<grid_for(n * nr, kBlock), kBlock, 0, s>>>(n, ns, nr, sp->states.p, flags.p, pos.p, d_kinds.p, pr,
                                                      A->sink_src.p, reinterpret_cast<VT*>(A->sink_val.p));
  }
  NC_CHECK_HIP(hipGetLastError());
  bool any_joint = false;
  for (int k : A->tv_kind) any_joint = any_joint || k == NC_PROP_JOINT;
  if (any_joint && n) {
    A->states.alloc((size_t)n * ns);
    A->state_conn.alloc((size_t)n * nr);
    NC_CHECK_HIP(hipMemcpyAsync(A->states.p, sp->states.p, sizeof(int32_t) * n * ns, hipMemcpyDeviceToDevice, s));
    NC_CHECK_HIP(hipMemcpyAsync(A->state_conn.p, sp->state_conn.p, sizeof(int32_t) * n * nr, hipMemcpyDeviceToDevice, s));
  }
  NC_CHECK_HIP(hipMemcpyAsync(dmax_h, d_dmax.p, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  NC_CHECK_HIP(hipStreamSynchronize(s));
  const unsigned long long dmax_bits = *dmax_h;
  A->dia_mask = 0;
  A->max_reach = 0;
  A->dia_off.assign(kd, 0);
  for (int k = 0; n && k < kd; ++k) {
    const int32_t lo = dia_h[2 * k], hi = dia_h[2 * k + 1];
    if (lo <= hi) A->max_reach = std::max<int64_t>(A->max_reach, std::max(std::abs((int64_t)lo), std::abs((int64_t)hi)));
    if (lo > hi || lo == hi) {  // no nonzero entry, or one offset throughout
      A->dia_mask |= 1u << k;
      A->dia_off[k] = lo > hi ? 0 : lo;
    }
  }
  double dmax;
  std::memcpy(&dmax, &dmax_bits, sizeof(dmax));
  A->norm1 = 2.0 * dmax;
  // nonzeros of the reference's matrices: merged TI matrix + one matrix per TV reaction
  std::vector<uint32_t> sinks_per_r(nr);
  for (int r = 0; r < nr; ++r) sinks_per_r[r] = (uint32_t)(A->sink_off_h[r + 1] - A->sink_off_h[r]);
  int64_t nnz = 0;
  if (g.ng) {
    nnz += n + (int64_t)counts[0];
    for (int r = 0; r < nr; ++r)
      if (m->kinds[r] == NC_PROP_TIME_INVARIANT) nnz += sinks_per_r[r];
  }
  for (int q = 0; q < tv.n; ++q) nnz += n + (int64_t)counts[2 + q] + sinks_per_r[tv.r[q]];
  A->nnz_ref = nnz;
  mf_setup(A, sp, m, g);
}

}  // namespace

double matrix_bytes_per_matvec(const nc_matrix_s* A) {
  const double vs = (double)A->vsize;
  if (A->mf_mode)  // computed values: v, out, the sink lists and the tables, plus the diagonal in mode 1
    return (double)A->n_sink_entries * (4.0 + vs) + (double)(A->nr + 1) * 8.0 + 2.0 * (double)A->n_rows * vs +
           8.0 * (double)A->mf_table_entries + (A->mf_mode == 1 ? 8.0 * (double)A->n : 0.0);
  const double n = (double)A->n;
  double b = n * 8.0                                  // diag
             + n * A->K * (4.0 + vs)                  // ELL slots
             - n * __builtin_popcount(A->dia_mask) * 4.0  // computed columns
             + n * A->tv_r.size() * (4.0 + 2.0 * vs)  // time-varying slots
             + (double)A->n_sink_entries * (4.0 + vs)  // sink lists
             + (double)(A->nr + 1) * 8.0;
  b += 2.0 * (double)A->n_rows * vs;  // read v, write out
  return b;
}

void matvec_harvest_timing(nc_matrix_s* A) {
  if (!A->ev_used) return;
  NC_CHECK_HIP(hipEventSynchronize(A->ev[2 * A->ev_used - 1]));
  for (size_t k = 0; k < A->ev_used; ++k) {
    float ms = 0.f;
    NC_CHECK_HIP(hipEventElapsedTime(&ms, A->ev[2 * k], A->ev[2 * k + 1]));
    A->timed_ms += ms;
    A->timed_count++;
  }
  A->ev_used = 0;
}

void matvec_launch(nc_matrix_s* A, double t, const void* v, void* out, bool accumulate, hipStream_t s,
                   bool out_fp64) {
  if (A->precision == NC_FP32)
    joint_refresh<float>(A, t);
  else
    joint_refresh<double>(A, t);
  if (s != A->stream.s) {
    // joint refresh runs on the matrix stream; order it before the launch on `s`
    hipEvent_t e;
    NC_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    NC_CHECK_HIP(hipEventRecord(e, A->stream.s));
    NC_CHECK_HIP(hipStreamWaitEvent(s, e, 0));
    NC_CHECK_HIP(hipEventDestroy(e));
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const bool timed = matvec_timing_slot(A, &e0, &e1);
  if (A->precision == NC_FP32)
    launch_matvec_t<float>(A, t, v, out, accumulate, s, out_fp64, e0, e1);
  else
    launch_matvec_t<double>(A, t, v, out, accumulate, s, out_fp64, e0, e1);
  NC_CHECK_HIP(hipGetLastError());
  if (timed) A->ev_used++;
  A->launches++;
}

bool matvec_rk_fusable(const nc_matrix_s* A) { return A->mf_mode == 0 && A->rows_per_thread == 2; }

void matvec_launch_rk(nc_matrix_s* A, double t, const void* v, double* out, hipStream_t s, const RkEpilogue& epi) {
  if (A->precision == NC_FP32)
    joint_refresh<float>(A, t);
  else
    joint_refresh<double>(A, t);
  if (s != A->stream.s) {
    hipEvent_t e;
    NC_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    NC_CHECK_HIP(hipEventRecord(e, A->stream.s));
    NC_CHECK_HIP(hipStreamWaitEvent(s, e, 0));
    NC_CHECK_HIP(hipEventDestroy(e));
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const bool timed = matvec_timing_slot(A, &e0, &e1);
  if (A->precision == NC_FP32)
    launch_matvec_t<float>(A, t, v, out, false, s, true, e0, e1, &epi);
  else
    launch_matvec_t<double>(A, t, v, out, false, s, true, e0, e1, &epi);
  NC_CHECK_HIP(hipGetLastError());
  if (timed) A->ev_used++;
  A->launches++;
}

int64_t matvec_err_parts(const nc_matrix_s* A) {
  const int64_t row_threads = (A->n + 1) / 2;  // 2 rows per thread
  return (row_threads + kBlock - 1) / kBlock + A->nr;
}

void matvec_launch_err(nc_matrix_s* A, double t, const void* v, double* out, hipStream_t s, const RkErrEpilogue& epi) {
  if (A->precision == NC_FP32)
    joint_refresh<float>(A, t);
  else
    joint_refresh<double>(A, t);
  if (s != A->stream.s) throw Error(NC_ERR_INTERNAL, "error-epilogue mat-vec off the matrix stream");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const bool timed = matvec_timing_slot(A, &e0, &e1);
  if (A->precision == NC_FP32)
    launch_matvec_t<float>(A, t, v, out, false, s, true, e0, e1, nullptr, nullptr, &epi);
  else
    launch_matvec_t<double>(A, t, v, out, false, s, true, e0, e1, nullptr, nullptr, &epi);
  NC_CHECK_HIP(hipGetLastError());
  if (timed) A->ev_used++;
  A->launches++;
}

bool matvec_dots_fusable(const nc_matrix_s* A) {
  if (A->mf_mode != 0 || A->rows_per_thread != 2 || (A->timing && A->timing_plain) || !A->tv_r.empty() || A->K < 1 ||
      A->K > 8)
    return false;
  const uint32_t all = (1u << A->K) - 1u;
  return A->dia_mask == 0u || (A->dia_mask & all) == all;
}

void matvec_launch_dots(nc_matrix_s* A, const void* v, double* out, hipStream_t s, const DualDots& d) {
  if (s != A->stream.s) throw Error(NC_ERR_INTERNAL, "projection mat-vec off the matrix stream");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const bool timed = matvec_timing_slot(A, &e0, &e1);
  if (A->precision == NC_FP32)
    launch_matvec_t<float>(A, 0.0, v, out, false, s, true, e0, e1, nullptr, &d);
  else
    launch_matvec_t<double>(A, 0.0, v, out, false, s, true, e0, e1, nullptr, &d);
  NC_CHECK_HIP(hipGetLastError());
  if (timed) A->ev_used++;
  A->launches++;
}

bool matvec_timing_slot(nc_matrix_s* A, hipEvent_t* e0, hipEvent_t* e1) {
  *e0 = *e1 = nullptr;
  if (!A->timing) return false;
  if (A->ev_used * 2 >= A->ev.size()) {
    if (A->ev.size() < 8192) {
      size_t old = A->ev.size();
      A->ev.resize(old ? old * 2 : 256);
      for (size_t k = old; k < A->ev.size(); ++k) NC_CHECK_HIP(hipEventCreate(&A->ev[k]));
    } else {
      matvec_harvest_timing(A);
    }
  }
  *e0 = A->ev[2 * A->ev_used];
  *e1 = A->ev[2 * A->ev_used + 1];
  return true;
}

}  // namespace nc

using namespace nc;

namespace {

template <typename VT>
__global__ void k_to_vt(const double* in, VT* out, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = (VT)in[i];
}

}  // namespace

// ---- fixed device scratch of the allocation-free mat-vec -------------------------------
// nc_matrix_matvec_noalloc stages its vectors here instead of allocating: the input
// (FP64), its FP32 copy for an FP32 matrix, and the FP64 output, each region aligned to
// 256 bytes. Every device the library runs on holds its own copy of the array until the
// process ends. In host code the array's name is not a device address: the address comes
// from hipGetSymbolAddress, taken once per device with that device current.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_dev[kMaxDevices] = {};  // device address of `scratch`, per device

unsigned char* scratch_on_current_device() {
  int dev = 0;
  NC_CHECK_HIP(hipGetDevice(&dev));
  NC_REQUIRE(dev >= 0 && dev < kMaxDevices, NC_ERR_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    NC_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    NC_REQUIRE(p, NC_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_dev[dev] = static_cast<unsigned char*>(p);
  }
  return g_scratch_dev[dev];
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Back-to-back launches of the integrators' mat-vec (stored-precision input, FP64 output).
double bench_matvec(nc_matrix_s* A, double t, int reps) {
  hipStream_t s = A->stream.s;
  const int64_t N = A->n_rows;
  DevBuf<unsigned char> v(N * A->vsize);
  DevBuf<double> o(N);
  NC_CHECK_HIP(hipMemsetAsync(v.p, 0, N * A->vsize, s));
  NC_CHECK_HIP(hipMemsetAsync(o.p, 0, N * sizeof(double), s));
  const bool was = A->timing;
  A->timing = false;
  matvec_launch(A, t, v.p, o.p, false, s, true);  // warm-up (and joint refresh)
  hipEvent_t e0, e1;
  NC_CHECK_HIP(hipEventCreate(&e0));
  NC_CHECK_HIP(hipEventCreate(&e1));
  NC_CHECK_HIP(hipEventRecord(e0, s));
  for (int k = 0; k < reps; ++k) matvec_launch(A, t, v.p, o.p, false, s, true);
  NC_CHECK_HIP(hipEventRecord(e1, s));
  NC_CHECK_HIP(hipEventSynchronize(e1));
  A->timing = was;
  float ms = 0.f;
  NC_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / reps;
}

// Mean duration of the mat-vec with the Krylov projection epilogue over nv basis vectors
// (device-resident zero vectors in the solver's padded layout), launched reps times back
// to back on the matrix's stream.
double bench_projection(nc_matrix_s* A, int nv, int reps) {
  const bool was = A->timing, was_plain = A->timing_plain;
  A->timing = false;
  A->timing_plain = false;
  struct Restore {
    nc_matrix_s* A;
    bool t, p;
    ~Restore() {
      A->timing = t;
      A->timing_plain = p;
    }
  } restore{A, was, was_plain};
  if (!matvec_dots_fusable(A))
    throw Error(NC_ERR_UNSUPPORTED, "this generator's mat-vec has no projection epilogue (it needs the stored "
                                    "generator, 2 rows per thread, 1..8 slots, no time-varying reactions)");
  NC_REQUIRE(nv >= 1 && nv <= 32, NC_ERR_ARG, "nv must be in 1..32");
  const int64_t nblk = (A->n + kDualRows - 1) / kDualRows;
  NC_REQUIRE(nblk >= 1 && nblk <= kDualMaxBlocks, NC_ERR_UNSUPPORTED, "state count outside the projection blocks' range");
  hipStream_t s = A->stream.s;
  const int64_t Np = (A->n_rows + 1023) / 1024 * 1024;
  DevBuf<unsigned char> V((size_t)nv * Np * A->vsize);
  DevBuf<double> o(Np), part((size_t)(2 * nv + 1) * nblk);
  NC_CHECK_HIP(hipMemsetAsync(V.p, 0, (size_t)nv * Np * A->vsize, s));
  NC_CHECK_HIP(hipMemsetAsync(o.p, 0, sizeof(double) * Np, s));
  const DualDots d{V.p, Np, nv, nullptr, part.p, nblk};
  const void* vlast = V.p + (size_t)(nv - 1) * Np * A->vsize;
  matvec_launch_dots(A, vlast, o.p, s, d);  // warm-up
  hipEvent_t e0, e1;
  NC_CHECK_HIP(hipEventCreate(&e0));
  NC_CHECK_HIP(hipEventCreate(&e1));
  NC_CHECK_HIP(hipEventRecord(e0, s));
  for (int k = 0; k < reps; ++k) matvec_launch_dots(A, vlast, o.p, s, d);
  NC_CHECK_HIP(hipEventRecord(e1, s));
  NC_CHECK_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  NC_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / reps;
}

// Streams n 16-byte words (reads only, so that no dirty lines are left behind) to evict
// whatever the caches hold; the xor is stored only if it hits an arbitrary constant, which
// keeps the loads alive.
__global__ void k_cache_flush(const uint4* p, int64_t n, unsigned* sink) {
  unsigned acc = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint4 w = p[i];
    acc ^= w.x ^ w.y ^ w.z ^ w.w;
  }
  if (acc == 0x9e3779b9u) sink[0] = acc;
}

// Median kernel time of the mat-vec with a cold matrix: before every timed launch a read
// of twice the 256 MB Infinity Cache evicts the matrix, as the basis streams of a Krylov
// step do between its mat-vecs, and the input vector is then rewritten, as the pass before
// a mat-vec in a solve writes it. Durations come from events stamped by the dispatch
// itself, so the flushes are not counted.
double bench_matvec_cold(nc_matrix_s* A, double t, int reps) {
  constexpr int64_t kFlushWords = (int64_t(512) << 20) / 16;
  hipStream_t s = A->stream.s;
  const int64_t N = A->n_rows;
  DevBuf<uint4> fl(kFlushWords);
  DevBuf<unsigned> sink(1);
  DevBuf<unsigned char> v(N * A->vsize), w(N * A->vsize);
  DevBuf<double> o(N);
  NC_CHECK_HIP(hipMemsetAsync(fl.p, 0, kFlushWords * 16, s));
  NC_CHECK_HIP(hipMemsetAsync(v.p, 0, N * A->vsize, s));
  NC_CHECK_HIP(hipMemsetAsync(w.p, 0, N * A->vsize, s));
  NC_CHECK_HIP(hipMemsetAsync(o.p, 0, N * sizeof(double), s));
  matvec_harvest_timing(A);
  const bool was = A->timing;
  const double ms0 = A->timed_ms;
  const int64_t n0 = A->timed_count;
  A->timing = false;
  matvec_launch(A, t, v.p, o.p, false, s, true);  // warm-up (and joint refresh)
  A->timing = true;
  A->timed_ms = 0.0;
  A->timed_count = 0;
  std::vector<float> each;
  for (int k = 0; k < reps; ++k) {
    // evict everything, then write the input vector: in a solve the mat-vec's input was
    // just written by the previous pass while the matrix comes from HBM
    unsigned char* in = (k & 1) ? w.p : v.p;
    k_cache_flush<<<4096, 256, 0, s>>>(fl.p, kFlushWords, sink.p);
    NC_CHECK_HIP(hipMemsetAsync(in, 0, N * A->vsize, s));
    matvec_launch(A, t, in, o.p, false, s, true);
    if (A->ev_used * 2 >= A->ev.size() || k == reps - 1) {
      // the launches' own durations, before the slots are reused
      NC_CHECK_HIP(hipEventSynchronize(A->ev[2 * A->ev_used - 1]));
      for (size_t q = 0; q < A->ev_used; ++q) {
        float ms = 0.f;
        NC_CHECK_HIP(hipEventElapsedTime(&ms, A->ev[2 * q], A->ev[2 * q + 1]));
        each.push_back(ms);
      }
      A->ev_used = 0;
    }
  }
  NC_CHECK_HIP(hipGetLastError());
  // median: single cold launches of ~10 us scatter by a few percent
  std::sort(each.begin(), each.end());
  const double ms = each.empty() ? 0.0 : (double)each[each.size() / 2];
  A->timing = was;
  A->timed_ms = ms0;
  A->timed_count = n0;
  NC_CHECK_HIP(hipStreamSynchronize(s));
  return ms;
}

}  // namespace

namespace nc {
void matrix_prepare(nc_matrix_s* A, double t) {
  if (A->precision == NC_FP32)
    joint_refresh<float>(A, t);
  else
    joint_refresh<double>(A, t);
}

void matrix_scales(const nc_matrix_s* A, double t, double* sink_scale, double* tv_scale) {
  const nc_model_s* m = A->model;
  for (int r = 0; r < A->nr; ++r) sink_scale[r] = m->kinds[r] == NC_PROP_SEPARABLE ? m->tfactor(r, t) : 1.0;
  for (size_t q = 0; q < A->tv_r.size(); ++q)
    tv_scale[q] = A->tv_kind[q] == NC_PROP_SEPARABLE ? m->tfactor(A->tv_r[q], t) : 1.0;
}

nc_matrix_s* matrix_create(nc_space_s* sp, nc_model_s* m, int precision) {
  NC_REQUIRE(precision == NC_FP64 || precision == NC_FP32, NC_ERR_ARG, "precision must be NC_FP64 or NC_FP32");
  NC_REQUIRE(sp->ns == m->ns && sp->nr == m->nr, NC_ERR_ARG, "model and state space sizes differ");
  NC_REQUIRE(sp->stoich == m->stoich, NC_ERR_ARG, "model and state space stoichiometry differ");
  auto A = new nc_matrix_s();
  try {
    A->precision = precision;
    A->vsize = precision == NC_FP32 ? 4 : 8;
    A->ns = sp->ns;
    A->nr = sp->nr;
    A->n = sp->n;
    A->n_rows = sp->n + sp->nr;
    A->model = m;
    NC_CHECK_HIP(hipStreamSynchronize(sp->stream.s));
    if (precision == NC_FP32)
      build_t<float>(A, sp, m);
    else
      build_t<double>(A, sp, m);
    NC_CHECK_HIP(hipStreamSynchronize(A->stream.s));
  } catch (...) {
    delete A;
    throw;
  }
  return A;
}
}  // namespace nc

extern "C" {

int nc_matrix_create(nc_space sp, nc_model m, int32_t precision, nc_matrix* out) {
  return guard([&] {
    NC_REQUIRE(sp && m && out, NC_ERR_ARG, "null argument");
    *out = nc::matrix_create(sp, m, precision);
  });
}

int nc_matrix_destroy(nc_matrix A) {
  return guard([&] { delete A; });
}

int nc_matrix_info(nc_matrix A, nc_matrix_info_t* info) {
  return guard([&] {
    NC_REQUIRE(A && info, NC_ERR_ARG, "null argument");
    info->n_rows = A->n_rows;
    info->n_states = A->n;
    info->n_sinks = A->nr;
    info->ell_width = A->K;
    info->n_tv = (int32_t)A->tv_r.size();
    info->precision = A->precision;
    info->nnz_reference = A->nnz_ref;
    info->n_sink_entries = A->n_sink_entries;
    info->bytes_per_matvec = matrix_bytes_per_matvec(A);
    const double sv = (double)A->vsize;
    info->csr_bytes_per_matvec =
        (double)A->nnz_ref * (sv + 4.0) + (double)(A->n_rows + 1) * 4.0 + 2.0 * (double)A->n_rows * sv;
    info->norm1 = A->norm1;
    info->rows_per_thread = A->rows_per_thread;
    info->computed_col_slots = __builtin_popcount(A->dia_mask);
    info->matrix_free = A->mf_mode;
    info->matrix_free_available = A->mf_ok ? (A->mf_multi ? 2 : 3) : 0;
  });
}

int nc_matrix_set_matrix_free(nc_matrix A, int32_t mode) {
  return guard([&] {
    NC_REQUIRE(A, NC_ERR_ARG, "null matrix");
    NC_REQUIRE(mode >= 0 && mode <= 2, NC_ERR_ARG, "mode must be 0 (stored), 1 (computed values) or 2 (matrix-free)");
    NC_REQUIRE(mode == 0 || A->mf_ok, NC_ERR_UNSUPPORTED,
               "no computed mat-vec for this generator: it needs a box state space in lattice order (species 0 "
               "fastest, at most 4 species) and time-invariant propensities that each read at most one species");
    NC_REQUIRE(mode != 1 || !A->mf_multi, NC_ERR_UNSUPPORTED,
               "computed values need every stoichiometry to hold one reaction and change one species");
    if (mode != A->mf_mode) A->launch_gen++;
    A->mf_mode = mode;
  });
}

int nc_matrix_set_rows_per_thread(nc_matrix A, int32_t rows) {
  return guard([&] {
    NC_REQUIRE(A, NC_ERR_ARG, "null matrix");
    NC_REQUIRE(rows == 1 || rows == 2 || rows == 4, NC_ERR_ARG, "rows per thread must be 1, 2 or 4");
    if (rows != A->rows_per_thread) A->launch_gen++;
    A->rows_per_thread = rows;
  });
}

int nc_matrix_autotune(nc_matrix A, int32_t reps, double* best_ms) {
  return guard([&] {
    NC_REQUIRE(A && reps > 0, NC_ERR_ARG, "bad arguments");
    double best = 1e300;
    const int old_r = A->rows_per_thread;
    int best_r = old_r;
    for (int r : {1, 2, 4}) {
      A->rows_per_thread = r;
      const double ms = bench_matvec_cold(A, 0.0, reps);
      if (ms < best) best = ms, best_r = r;
    }
    A->rows_per_thread = best_r;
    if (best_r != old_r) A->launch_gen++;
    if (best_ms) *best_ms = best;
  });
}

int nc_matrix_matvec(nc_matrix A, double t, const double* v, double* out, int32_t accumulate) {
  return guard([&] {
    NC_REQUIRE(A && v && out, NC_ERR_ARG, "null argument");
    hipStream_t s = A->stream.s;
    const int64_t N = A->n_rows;
    DevBuf<double> dv(N), dout(N);
    NC_CHECK_HIP(hipMemcpyAsync(dv.p, v, sizeof(double) * N, hipMemcpyHostToDevice, s));
    if (accumulate) NC_CHECK_HIP(hipMemcpyAsync(dout.p, out, sizeof(double) * N, hipMemcpyHostToDevice, s));
    if (A->precision == NC_FP32) {
      // FP32 storage rounds the matrix values and the input; the kernel accumulates and
      // writes (or adds into) the output in FP64, as inside the integrators
      DevBuf<float> vv(N);
      k_to_vt<float><<<grid_for(N, 256), 256, 0, s>>>(dv.p, vv.p, N);
      matvec_launch(A, t, vv.p, dout.p, accumulate != 0, s, true);
    } else {
      matvec_launch(A, t, dv.p, dout.p, accumulate != 0, s, true);
    }
    NC_CHECK_HIP(hipGetLastError());
    NC_CHECK_HIP(hipMemcpyAsync(out, dout.p, sizeof(double) * N, hipMemcpyDeviceToHost, s));
    NC_CHECK_HIP(hipStreamSynchronize(s));
    A->ev_used = 0;
  });
}

int nc_matrix_matvec_noalloc(nc_matrix A, double t, const double* v, double* out, int32_t accumulate) {
  return guard([&] {
    NC_REQUIRE(A && v && out, NC_ERR_ARG, "null argument");
    const int64_t N = A->n_rows;
    const bool fp32 = A->precision == NC_FP32;
    const size_t b64 = align256(sizeof(double) * (size_t)N), b32 = align256(sizeof(float) * (size_t)N);
    const size_t need = 2 * b64 + (fp32 ? b32 : 0);
    NC_REQUIRE(need <= kNoAllocScratchBytes, NC_ERR_ARG,
               "nc_matrix_matvec_noalloc: " + std::to_string(N) + " rows need " + std::to_string(need) +
                   " bytes of scratch, more than its fixed " + std::to_string(kNoAllocScratchBytes) +
                   " (about 2.5M rows at FP32 storage, 3.1M at FP64); use nc_matrix_matvec");
    unsigned char* base = scratch_on_current_device();
    double* dv = reinterpret_cast<double*>(base);
    double* dout = reinterpret_cast<double*>(base + b64);
    hipStream_t s = A->stream.s;
    NC_CHECK_HIP(hipMemcpyAsync(dv, v, sizeof(double) * N, hipMemcpyHostToDevice, s));
    if (accumulate) NC_CHECK_HIP(hipMemcpyAsync(dout, out, sizeof(double) * N, hipMemcpyHostToDevice, s));
    if (fp32) {
      float* vf = reinterpret_cast<float*>(base + 2 * b64);
      k_to_vt<float><<<grid_for(N, 256), 256, 0, s>>>(dv, vf, N);
      matvec_launch(A, t, vf, dout, accumulate != 0, s, true);
    } else {
      matvec_launch(A, t, dv, dout, accumulate != 0, s, true);
    }
    NC_CHECK_HIP(hipGetLastError());
    NC_CHECK_HIP(hipMemcpyAsync(out, dout, sizeof(double) * N, hipMemcpyDeviceToHost, s));
    NC_CHECK_HIP(hipStreamSynchronize(s));
    A->ev_used = 0;
  });
}

int nc_matrix_export(nc_matrix A, double t, int64_t* nnz, int64_t* rows, int64_t* cols, double* vals) {
  return guard([&] {
    NC_REQUIRE(A && nnz, NC_ERR_ARG, "null argument");
    const int64_t n = A->n, ld = A->ld;
    const int ntv = (int)A->tv_r.size();
    const int64_t total = n * (1 + A->K + 2 * ntv) + A->n_sink_entries;
    *nnz = total;
    if (!rows) return;
    NC_REQUIRE(cols && vals, NC_ERR_ARG, "null output arrays");
    if (A->precision == NC_FP32)
      joint_refresh<float>(A, t);
    else
      joint_refresh<double>(A, t);
    NC_CHECK_HIP(hipStreamSynchronize(A->stream.s));
    auto fetch = [&](const unsigned char* src, int64_t count) {
      std::vector<double> out(count);
      if (!count) return out;
      if (A->vsize == 8) {
        NC_CHECK_HIP(hipMemcpy(out.data(), src, 8 * count, hipMemcpyDeviceToHost));
      } else {
        std::vector<float> f(count);
        NC_CHECK_HIP(hipMemcpy(f.data(), src, 4 * count, hipMemcpyDeviceToHost));
        for (int64_t k = 0; k < count; ++k) out[k] = f[k];
      }
      return out;
    };
    std::vector<double> diag(ld);
    if (ld) NC_CHECK_HIP(hipMemcpy(diag.data(), A->diag.p, 8 * ld, hipMemcpyDeviceToHost));
    std::vector<int32_t> col((size_t)A->K * ld), tvc((size_t)ntv * ld), src(A->n_sink_entries);
    if (!col.empty()) NC_CHECK_HIP(hipMemcpy(col.data(), A->col.p, 4 * col.size(), hipMemcpyDeviceToHost));
    if (!tvc.empty()) NC_CHECK_HIP(hipMemcpy(tvc.data(), A->tvc.p, 4 * tvc.size(), hipMemcpyDeviceToHost));
    if (!src.empty()) NC_CHECK_HIP(hipMemcpy(src.data(), A->sink_src.p, 4 * src.size(), hipMemcpyDeviceToHost));
    auto val = fetch(A->val.p, (int64_t)A->K * ld);
    auto tvd = fetch(A->tvd.p, (int64_t)ntv * ld);
    auto tvv = fetch(A->tvv.p, (int64_t)ntv * ld);
    auto sv = fetch(A->sink_val.p, A->n_sink_entries);
    const nc_model_s* m = A->model;
    int64_t k = 0;
    for (int64_t i = 0; i < n; ++i) rows[k] = i, cols[k] = i, vals[k] = diag[i], ++k;
    for (int s = 0; s < A->K; ++s)
      for (int64_t i = 0; i < n; ++i) rows[k] = i, cols[k] = col[s * ld + i], vals[k] = val[s * ld + i], ++k;
    for (int q = 0; q < ntv; ++q) {
      const double sc = A->tv_kind[q] == NC_PROP_SEPARABLE ? m->tfactor(A->tv_r[q], t) : 1.0;
      for (int64_t i = 0; i < n; ++i) rows[k] = i, cols[k] = i, vals[k] = sc * tvd[q * ld + i], ++k;
      for (int64_t i = 0; i < n; ++i) rows[k] = i, cols[k] = tvc[q * ld + i], vals[k] = sc * tvv[q * ld + i], ++k;
    }
    for (int r = 0; r < A->nr; ++r) {
      const double sc = m->kinds[r] == NC_PROP_SEPARABLE ? m->tfactor(r, t) : 1.0;
      for (int64_t e = A->sink_off_h[r]; e < A->sink_off_h[r + 1]; ++e)
        rows[k] = n + r, cols[k] = src[e], vals[k] = sc * sv[e], ++k;
    }
  });
}

int nc_matrix_bench_matvec(nc_matrix A, double t, int32_t reps, double* mean_ms) {
  return guard([&] {
    NC_REQUIRE(A && mean_ms && reps > 0, NC_ERR_ARG, "bad arguments");
    *mean_ms = bench_matvec(A, t, reps);
  });
}

int nc_matrix_bench_projection(nc_matrix A, int32_t nv, int32_t reps, double* mean_ms) {
  return guard([&] {
    NC_REQUIRE(A && mean_ms && reps > 0, NC_ERR_ARG, "bad arguments");
    *mean_ms = bench_projection(A, nv, reps);
  });
}

}  // extern "C"
