// Device-resident time integration of the FSP ODE du/dt = A(t) u.
//
// Replaces the integrator loop of NumCME.jl's solve (src/transientcme/sparse/fspsolve.jl:
// 10-41, a DifferentialEquations.jl/CVODE call around matvec!) with:
//
//  * Krylov expv for time-invariant generators: Arnoldi by classical Gram-Schmidt whose
//    coefficients solve the Gram system of the stored basis (one projection pass and one
//    update pass per column, orthogonal to FP64 rounding even with an FP32 basis), FP64
//    dot products and coefficients, basis vectors stored in the matrix precision; the
//    (m+2)x(m+2) Hessenberg exponential, the local error estimate and the step-size
//    control run on the host, following R. B. Sidje, "Expokit: a software package for
//    computing matrix exponentials", ACM TOMS 24(1), 1998 (algorithm of its EXPV
//    routine). One host synchronisation per Krylov step.
//  * Dormand-Prince 5(4) (J. R. Dormand & P. J. Prince, J. Comput. Appl. Math. 6, 1980)
//    with FSAL, the Hairer-Norsett-Wanner RMS error norm computed on the device in FP64,
//    and output times hit exactly by step clamping.
//
// Vectors: the master state u is FP64; mat-vec inputs are stored in the matrix precision.
#include <algorithm>
#include <cmath>
#include <limits>
#include <type_traits>

#include "nc_dual.h"
#include "nc_fspvec.h"
#include "nc_solver.h"

namespace nc {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxRedBlocks = 2048;
constexpr int kPartStrideMax = 65;  // partial sums per block: up to 2 x 32 dots + a norm

// FIN_FOLD: the finish kernel only folds the partials into a.tot (the next pass, an
// update with arn_pre set, applies the Arnoldi finishing step itself)
enum Finish : int { FIN_NONE = 0, FIN_LAST, FIN_BETA, FIN_SUMS, FIN_FOLD };
constexpr double kBreakdownTol = 1e-7;   // Expokit's happy-breakdown threshold
// threads of each fold workgroup: one partial per thread at the 1,024-workgroup grid (a
// thread looping over several would wait for each load in turn)
constexpr int kFinBlock = 1024;
constexpr unsigned kVecBlocks = 512;  // grid of the vector passes
constexpr int kLoadChunk = 8;  // basis vectors whose loads a vector pass issues together
constexpr int kJacobiSweeps = 3;  // G h = a with G = I + O(1e-7): the error falls 1e-6 per sweep

// kbuf layout (doubles)
struct KLayout {
  int m;
  __host__ __device__ int H(int i, int j) const { return i * (m + 2) + j; }
  __host__ __device__ int h(int i) const { return (m + 2) * (m + 2) + i; }
  __host__ __device__ int sc(int i) const { return (m + 2) * (m + 2) + (m + 2) + i; }
  __host__ __device__ int misc(int i) const { return (m + 2) * (m + 2) + 2 * (m + 2) + i; }
  // sink rows of basis vector j (j <= m) and of the state (j = m+1), R sinks each
  __host__ __device__ int sinkv(int j, int r, int R) const { return misc(8) + j * R + r; }
  __host__ __device__ int size() const { return (m + 2) * (m + 2) + 2 * (m + 2) + 8; }
  __host__ __device__ int size_with_sinks(int R) const { return size() + (m + 2) * R; }
};
enum { MISC_AVN = 0, MISC_BETA = 1 };

template <typename BT>
struct VecArgs {
  int64_t N;
  int nv;
  const BT* V;            // nv x N basis (raw)
  const double* sc;       // per-vector scale (device, may be null = 1)
  const double* coef;     // per-vector coefficient (device, may be null)
  double csign;           // multiplies coef
  const double* xin;      // FP64 input
  const double* xscale;   // device scalar multiplying xin (may be null = 1)
  double* xout;           // FP64 outputs (nullable)
  void* vout;             // VT output (nullable)
  int vout_fp32;          // 1: vout is float, 0: double
  int mode;               // VM_*
  double* partials;
  int finish;
  double* kb;
  double* gram;           // (m+1) x (m+1) Gram matrix of the scaled basis (arn_pre)
  double* tot;            // kPartStrideMax folded totals (finish)
  int arn_pre;            // VM_UPDATE: apply column j's Arnoldi finishing step from tot first
  unsigned* counter;      // arrival counter of the fold's workgroups, zero between passes
  KLayout L;
  int j;
  double btol;
  int64_t tail0, tail1;   // fold: rows [tail0, tail1) of V and xin (the sink rows) added to the
                          // block sums of the 512-row projection passes
};

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Partial sums are stored value-major, partials[k * nblocks + b], so the finish kernel's
// loads of one value across all blocks are contiguous.
//
// x = xscale*xin + csign * sum_i coef[i]*sc[i]*V_i, with optional FP64 / VT stores of x,
// and per-block partial sums of what the pass measures:
//   VM_DOTS    x = xscale*xin; d_i = sc[i] * sum_n V_i[n] x[n]        (nv dots)
//   VM_UPDATE  the update, no dots
//   VM_DUAL    x = xin; p_i = sum_n V_i[n] x[n] and g_i = sum_n V_i[n] V_last[n]
//              with V_last = V_{nv-1}                                  (2 nv dots)
// and always the norm sum_n x[n]^2 last. Each thread owns 4 consecutive rows and reads
// every stream with one 16-byte (FP32) or 32-byte (FP64) vector load per vector; N is a
// multiple of 4 and all vectors are zero beyond the true length, so no tail handling is
// needed.
enum VecMode : int { VM_DOTS = 0, VM_UPDATE = 1, VM_DUAL = 2 };

template <typename T, int R>
struct alignas(sizeof(T) * R) Pack {
  T v[R];
};

// Finishing step of Arnoldi column j from the folded totals tot (V_i.q at i, V_i.V_j at
// nv + i, nv = j + 1), done by every workgroup of the update pass that follows instead of
// by one workgroup in a launch of its own: |V_j| (H(j, j-1), sc_j), the Gram column of the
// scaled basis, and h from G h = a, a_i = sc_i sc_j V_i.q, by Jacobi sweeps (G is the
// identity up to the FP32 rounding of the stored basis). Leaves the update coefficients
// -h_i sc_i in cs[0..nv) and sc_j in *xs; workgroup 0 also stores H(., j), h, sc_j and the
// Gram column for the later columns (no other workgroup reads those entries). kBlock
// threads: the sweeps give each row i eight lanes.
__device__ void arnoldi_prologue(const KLayout& L, double* kb, double* gram, int j, double btol, const double* tot,
                                 double* cs, double* xs) {
  static_assert(kBlock == 256, "arnoldi_prologue assumes 256 threads");
  const int nv = j + 1, ld = L.m + 1;
  __shared__ double p_scj, p_nj, p_a[32], p_h[32], p_sc[32], p_tot[kPartStrideMax], p_G[32][33];
  const int tid = threadIdx.x;
  // every global read in one round trip: the previous Gram block, the scales, the totals
  for (int e = tid; e < j * j; e += kBlock) p_G[e / j][e % j] = gram[(e / j) * ld + e % j];
  if (tid < j || (j == 0 && tid == 0)) p_sc[tid] = kb[L.sc(tid)];  // sc(j), j >= 1, is set below
  if (tid < 2 * nv) p_tot[tid] = tot[tid];
  __syncthreads();
  if (tid == 0) {
    double nj = 0.0, scj = p_sc[0];
    if (j >= 1) {
      nj = sqrt(p_tot[nv + j]);
      scj = nj > btol ? 1.0 / nj : 0.0;
    }
    p_nj = nj;
    p_scj = scj;
  }
  __syncthreads();
  const double scj = p_scj;
  if (tid < nv) {
    const int r = tid;
    const double scr = r == j ? scj : p_sc[r];
    double grj = scr * scj * p_tot[nv + r];
    if (r == j && scj == 0.0) grj = 1.0;  // a vanished vector: keep G invertible
    p_G[r][j] = grj;
    p_G[j][r] = grj;
    const double ar = scr * scj * p_tot[r];
    p_a[r] = ar;
    p_h[r] = ar;
  }
  if (tid == 0) p_sc[j] = scj;
  __syncthreads();
  const int i = tid >> 3, kq = tid & 7;
  for (int sweep = 0; sweep < kJacobiSweeps; ++sweep) {
    double t = 0.0;
    if (i < nv)
      for (int k = kq; k < nv; k += 8) t += (k == i ? p_G[i][k] - 1.0 : p_G[i][k]) * p_h[k];
    for (int off = 4; off > 0; off >>= 1) t += __shfl_down(t, off, 8);
    __syncthreads();
    if (kq == 0 && i < nv) p_h[i] = p_a[i] - t;
    __syncthreads();
  }
  if (tid < nv) cs[tid] = -p_h[tid] * p_sc[tid];
  if (tid == 0) *xs = scj;
  if (blockIdx.x == 0) {
    if (tid == 0 && j >= 1) {
      kb[L.H(j, j - 1)] = p_nj;
      kb[L.sc(j)] = scj;
    }
    if (tid < nv) {
      kb[L.H(tid, j)] = p_h[tid];
      kb[L.h(tid)] = p_h[tid];
      gram[tid * ld + j] = p_G[tid][j];
      gram[j * ld + tid] = p_G[j][tid];
    }
  }
}

template <typename BT, int MAXV, int MODE>
__global__ __launch_bounds__(kBlock) void k_vec(VecArgs<BT> a) {
  constexpr int MV = MAXV > 0 ? MAXV : 1;
  constexpr bool DOTS = MODE != VM_UPDATE;
  constexpr bool UPD = MODE == VM_UPDATE;
  constexpr bool DUAL = MODE == VM_DUAL;
  constexpr int R = 4;
  const int nv = a.nv;
  // per-vector coefficients live in LDS (uniform broadcast reads)
  __shared__ double s_cs[MV], s_sc[MV], s_xs;
  if (threadIdx.x < MV) {
    const int i = threadIdx.x;
    const double sc = (i < nv) ? (a.sc ? a.sc[i] : 1.0) : 0.0;
    s_sc[i] = sc;
    s_cs[i] = (i < nv && a.coef) ? a.csign * a.coef[i] * sc : 0.0;
  }
  if (threadIdx.x == 0) s_xs = a.xscale ? *a.xscale : 1.0;
  __syncthreads();
  const double xs = s_xs;
  double dacc[DOTS ? MV : 1], gacc[DUAL ? MV : 1];
#pragma unroll
  for (int i = 0; i < (DOTS ? MV : 1); ++i) dacc[i] = 0.0;
#pragma unroll
  for (int i = 0; i < (DUAL ? MV : 1); ++i) gacc[i] = 0.0;
  double nacc = 0.0;
  const int64_t groups = a.N / R;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x; g < groups; g += stride) {
    const int64_t n0 = g * R;
    double x[R];
    const Pack<double, R> xi = *reinterpret_cast<const Pack<double, R>*>(a.xin + n0);
    Pack<BT, R> vl;
    if constexpr (DUAL) vl = *reinterpret_cast<const Pack<BT, R>*>(a.V + (int64_t)(nv - 1) * a.N + n0);
    if constexpr (MAXV == 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) x[r] = xs * xi.v[r];
    }
    // The basis is read in chunks of kLoadChunk vectors whose loads are issued together
    // (unconditional, index clamped to nv - 1: a guarded load becomes a branch that
    // waits for it) and kept ahead of their arithmetic by a scheduling barrier; left to
    // itself the scheduler reuses one register set and waits after every load, which
    // streams the basis one vector at a time. Vectors beyond nv carry a zero coefficient
    // (update) or land in dots that are never stored.
#pragma unroll
    for (int i0 = 0; i0 < MAXV; i0 += kLoadChunk) {
      Pack<BT, R> vv[kLoadChunk];
#pragma unroll
      for (int u = 0; u < kLoadChunk; ++u)
        if (i0 + u < MAXV)
          vv[u] = *reinterpret_cast<const Pack<BT, R>*>(a.V + (int64_t)min(i0 + u, nv - 1) * a.N + n0);
      __builtin_amdgcn_sched_barrier(0);
      if (i0 == 0) {
#pragma unroll
        for (int r = 0; r < R; ++r) x[r] = xs * xi.v[r];
      }
#pragma unroll
      for (int u = 0; u < kLoadChunk; ++u) {
        const int i = i0 + u;
        if (i >= MAXV) break;
        if constexpr (UPD) {
#pragma unroll
          for (int r = 0; r < R; ++r) x[r] += s_cs[i] * (double)vv[u].v[r];
        } else {
#pragma unroll
          for (int r = 0; r < R; ++r) dacc[i] += (double)vv[u].v[r] * x[r];
        }
        if constexpr (DUAL) {
#pragma unroll
          for (int r = 0; r < R; ++r) gacc[i] += (double)vv[u].v[r] * (double)vl.v[r];
        }
      }
    }
    if (a.xout) {
      Pack<double, R> o;
#pragma unroll
      for (int r = 0; r < R; ++r) o.v[r] = x[r];
      *reinterpret_cast<Pack<double, R>*>(a.xout + n0) = o;
    }
    if (a.vout) {
      if (a.vout_fp32) {
        Pack<float, R> o;
#pragma unroll
        for (int r = 0; r < R; ++r) o.v[r] = (float)x[r];
        *reinterpret_cast<Pack<float, R>*>(reinterpret_cast<float*>(a.vout) + n0) = o;
      } else {
        Pack<double, R> o;
#pragma unroll
        for (int r = 0; r < R; ++r) o.v[r] = x[r];
        *reinterpret_cast<Pack<double, R>*>(reinterpret_cast<double*>(a.vout) + n0) = o;
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) nacc += x[r] * x[r];
  }
  if (a.finish == FIN_NONE) return;
  // block reduction of the dots (scaled by sc[i] here) and the norm
  __shared__ double red[2 * MV + 1][kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ndots = DOTS ? (DUAL ? 2 * nv : nv) : 0;
  if constexpr (DOTS) {
#pragma unroll
    for (int i = 0; i < MAXV; ++i) {
      if (i < nv) {
        double v = wave_sum(dacc[i]) * s_sc[i];
        if (lane == 0) red[i][wave] = v;
        if constexpr (DUAL) {
          double w = wave_sum(gacc[i]);
          if (lane == 0) red[nv + i][wave] = w;
        }
      }
    }
  }
  {
    double v = wave_sum(nacc);
    if (lane == 0) red[ndots][wave] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x <= ndots) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) s += red[threadIdx.x][w];
    a.partials[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
  }
  // k_vec_finish folds the partials in the next launch: the kernel boundary makes them
  // visible across XCDs without a per-block L2 writeback fence
}

// The Arnoldi update V_{j+1} = sc_j q - sum_i h_i sc_i V_i with the column's finishing step
// (arnoldi_prologue) in every workgroup. Each thread issues its first q rows and its first
// kLoadChunk basis rows before the prologue (its barriers wait for LDS only, so the loads
// stay in flight while it runs), which hides the prologue behind the streams.
template <typename BT, int MAXV>
__global__ __launch_bounds__(kBlock) void k_arn_update(VecArgs<BT> a) {
  constexpr int R = 4, C0 = MAXV < kLoadChunk ? MAXV : kLoadChunk;
  const int nv = a.nv;
  const int64_t groups = a.N / R, stride = (int64_t)gridDim.x * kBlock;
  int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t gc = g < groups ? g : groups - 1;  // clamped: the loads are unconditional
  Pack<double, R> xi = *reinterpret_cast<const Pack<double, R>*>(a.xin + gc * R);
  Pack<BT, R> v0[C0];
#pragma unroll
  for (int u = 0; u < C0; ++u) v0[u] = *reinterpret_cast<const Pack<BT, R>*>(a.V + (int64_t)min(u, nv - 1) * a.N + gc * R);
  __shared__ double s_cs[MAXV], s_xs;
  if (threadIdx.x < MAXV) s_cs[threadIdx.x] = 0.0;
  __syncthreads();
  arnoldi_prologue(a.L, a.kb, a.gram, a.j, a.btol, a.tot, s_cs, &s_xs);
  __syncthreads();
  const double xs = s_xs;
  BT* vout = reinterpret_cast<BT*>(a.vout);
  for (; g < groups; g += stride) {
    const int64_t n0 = g * R;
    if (g != gc) {  // a later grid-stride row group (not prefetched)
      xi = *reinterpret_cast<const Pack<double, R>*>(a.xin + n0);
#pragma unroll
      for (int u = 0; u < C0; ++u) v0[u] = *reinterpret_cast<const Pack<BT, R>*>(a.V + (int64_t)min(u, nv - 1) * a.N + n0);
    }
    double x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = xs * xi.v[r];
#pragma unroll
    for (int u = 0; u < C0; ++u) {
#pragma unroll
      for (int r = 0; r < R; ++r) x[r] += s_cs[u] * (double)v0[u].v[r];
    }
#pragma unroll
    for (int i0 = C0; i0 < MAXV; i0 += kLoadChunk) {
      Pack<BT, R> vv[kLoadChunk];
#pragma unroll
      for (int u = 0; u < kLoadChunk; ++u)
        if (i0 + u < MAXV)
          vv[u] = *reinterpret_cast<const Pack<BT, R>*>(a.V + (int64_t)min(i0 + u, nv - 1) * a.N + n0);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < kLoadChunk; ++u) {
        if (i0 + u >= MAXV) break;
#pragma unroll
        for (int r = 0; r < R; ++r) x[r] += s_cs[i0 + u] * (double)vv[u].v[r];
      }
    }
    Pack<BT, R> o;
#pragma unroll
    for (int r = 0; r < R; ++r) o.v[r] = (BT)x[r];
    *reinterpret_cast<Pack<BT, R>*>(vout + n0) = o;
  }
}

// VM_DUAL with the basis split across the workgroup's waves (wave w owns V_w, V_{w+4},
// ...). Every wave streams the same rows of x and V_last, which the other waves' reads
// keep in L2, and only its own basis vectors, so a thread holds 2 ceil(MAXV / 4)
// accumulators instead of 2 MAXV and the pass keeps its occupancy at high column counts.
// Same partial-sum layout as k_vec (each value is written by the wave that owns it).
template <typename BT, int MAXV>
__global__ __launch_bounds__(kBlock) void k_vec_dual(VecArgs<BT> a) {
  constexpr int W = kBlock / 64, MVW = (MAXV + W - 1) / W, R = 4;
  const int nv = a.nv;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const double xs = a.xscale ? *a.xscale : 1.0;
  double dacc[MVW], gacc[MVW], nacc = 0.0;
#pragma unroll
  for (int k = 0; k < MVW; ++k) dacc[k] = gacc[k] = 0.0;
  const BT* Vl = a.V + (int64_t)(nv - 1) * a.N;
  const int64_t groups = a.N / R;
  const int64_t stride = (int64_t)gridDim.x * 64;
  for (int64_t g = (int64_t)blockIdx.x * 64 + lane; g < groups; g += stride) {
    const int64_t n0 = g * R;
    const Pack<double, R> xi = *reinterpret_cast<const Pack<double, R>*>(a.xin + n0);
    const Pack<BT, R> vl = *reinterpret_cast<const Pack<BT, R>*>(Vl + n0);
    // all of this wave's basis loads issued together (unconditional, clamped; see k_vec)
    Pack<BT, R> vv[MVW];
#pragma unroll
    for (int k = 0; k < MVW; ++k)
      vv[k] = *reinterpret_cast<const Pack<BT, R>*>(a.V + (int64_t)min(k * W + wave, nv - 1) * a.N + n0);
    __builtin_amdgcn_sched_barrier(0);
    double x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = xs * xi.v[r];
#pragma unroll
    for (int k = 0; k < MVW; ++k) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        dacc[k] += (double)vv[k].v[r] * x[r];
        gacc[k] += (double)vv[k].v[r] * (double)vl.v[r];
      }
    }
    if (wave == 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) nacc += x[r] * x[r];
    }
  }
  const int64_t nb = gridDim.x;
#pragma unroll
  for (int k = 0; k < MVW; ++k) {
    const int i = k * W + wave;
    if (i < nv) {
      const double d = wave_sum(dacc[k]) * (a.sc ? a.sc[i] : 1.0);
      const double gm = wave_sum(gacc[k]);
      if (lane == 0) {
        a.partials[(int64_t)i * nb + blockIdx.x] = d;
        a.partials[(int64_t)(nv + i) * nb + blockIdx.x] = gm;
      }
    }
  }
  if (wave == 0) {
    const double v = wave_sum(nacc);
    if (lane == 0) a.partials[(int64_t)(2 * nv) * nb + blockIdx.x] = v;
  }
}

// Fold of the per-block partials, then the finishing step. One workgroup per value:
// workgroup k sums value k over the pass's blocks in a fixed order (thread t takes blocks
// t, t + 1024, ..., then a fixed shuffle tree), so the totals are deterministic, and
// stores it. The workgroup that arrives last at a device-scope counter (release fence
// before the increment, acquire fence after it) reads every total, re-arms the counter
// for the next pass and applies the finishing step. The values are folded by nvals
// compute units at once instead of one.
template <typename BT>
__global__ __launch_bounds__(kFinBlock) void k_vec_finish(VecArgs<BT> a, unsigned nblocks, int nvals) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __shared__ double tot[kPartStrideMax];
  __shared__ double wsum[kFinBlock / 64];
  __shared__ int s_last;
  {
    const int k = blockIdx.x;
    // thread t takes blocks t, t + 1024, t + 2048, t + 3072 (nblocks <= 4096), loaded
    // together and added in that order
    double pv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const unsigned b = threadIdx.x + u * kFinBlock;
      const double t = a.partials[(int64_t)k * nblocks + min(b, nblocks - 1)];
      pv[u] = b < nblocks ? t : 0.0;
    }
    double v = 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) v += pv[u];
    v = wave_sum(v);
    if (lane == 0) wsum[wave] = v;
    // the rows past the blocks (sink rows of the 512-row projection passes), by wave 0 in
    // row order; value k is V_k.x (times sc_k), V_{k-nv}.V_{nv-1} or x.x as in the blocks
    double tl = 0.0;
    if (wave == 0 && a.tail1 > a.tail0) {
      const int nv = a.nv;
      const BT* Vl = a.V + (int64_t)(nv - 1) * a.N;
      for (int64_t o = a.tail0 + lane; o < a.tail1; o += 64) {
        const double xo = a.xin[o];
        if (k < nv)
          tl = __fma_rn((double)a.V[(int64_t)k * a.N + o], xo, tl);
        else if (k < 2 * nv)
          tl = __fma_rn((double)a.V[(int64_t)(k - nv) * a.N + o], (double)Vl[o], tl);
        else
          tl = __fma_rn(xo, xo, tl);
      }
      tl = wave_sum(tl);
      if (k < nv && a.sc) tl *= a.sc[k];
    }
    __syncthreads();
    if (a.finish == FIN_FOLD) {
      // the totals are read by the next launch: its kernel boundary makes them visible
      if (threadIdx.x == 0) {
        double sum = 0.0;
        for (int w = 0; w < kFinBlock / 64; ++w) sum += wsum[w];
        a.tot[k] = sum + tl;
      }
      return;
    }
    if (threadIdx.x == 0) {
      double sum = 0.0;
      for (int w = 0; w < kFinBlock / 64; ++w) sum += wsum[w];
      sum += tl;
      __hip_atomic_store(reinterpret_cast<unsigned long long*>(a.tot) + k, (unsigned long long)__double_as_longlong(sum),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __threadfence();
      const unsigned prev = atomicAdd(a.counter, 1u);
      const int last = prev + 1u == gridDim.x;
      if (last) __threadfence();
      s_last = last;
    }
    __syncthreads();
    if (!s_last) return;
    if ((int)threadIdx.x < nvals)
      tot[threadIdx.x] = __longlong_as_double((long long)__hip_atomic_load(
          reinterpret_cast<unsigned long long*>(a.tot) + threadIdx.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    if (threadIdx.x == 0) __hip_atomic_store(a.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  const KLayout& L = a.L;
  double* kb = a.kb;
  if (threadIdx.x == 0) {
    const double nrm2 = tot[nvals - 1];
    switch (a.finish) {
      case FIN_LAST: {
        // VM_DUAL over V_m alone with x = A V_m: norm of V_m and |A v_m| for the error
        // estimate (avnorm of Expokit)
        const int m = L.m;
        const double nm = sqrt(tot[1]);
        const double scm = nm > a.btol ? 1.0 / nm : 0.0;
        kb[L.H(m, m - 1)] = nm;
        kb[L.sc(m)] = scm;
        kb[L.misc(MISC_AVN)] = scm * sqrt(nrm2);
      } break;
      case FIN_BETA: {
        const double b = sqrt(nrm2);
        kb[L.misc(MISC_BETA)] = b;
        kb[L.sc(0)] = b > 0.0 ? 1.0 / b : 0.0;
      } break;
      case FIN_SUMS:
        for (int i = 0; i < nvals; ++i) kb[i] = tot[i];
        break;
      default: break;
    }
  }
}

// The projection sums of Arnoldi column nv - 1 over 512-row blocks from the stored
// q = A V_{nv-1} (nc_dual.h): the same block sums the mat-vec's projection epilogue forms,
// for generators whose mat-vec has none. One workgroup per block.
template <typename BT>
__global__ __launch_bounds__(kBlock) void k_dual_blocks(const BT* __restrict__ V, int64_t ld, int nv,
                                                        const double* __restrict__ q, int64_t n, double* part) {
  const int64_t blk = blockIdx.x, base = blk * kDualRows;
  const int lane = threadIdx.x & 63;
  const BT* Vl = V + (int64_t)(nv - 1) * ld;
  double x[8], y[8];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t r = base + 256 * h + 4 * lane;
    const Pack<double, 4> qx = *reinterpret_cast<const Pack<double, 4>*>(q + r);
    const Pack<BT, 4> yl = *reinterpret_cast<const Pack<BT, 4>*>(Vl + r);
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const bool live = r + rr < n;
      x[4 * h + rr] = live ? qx.v[rr] : 0.0;
      y[4 * h + rr] = live ? (double)yl.v[rr] : 0.0;
    }
  }
  dual_block<BT>(V, ld, nv, base, x, y, nullptr, part, gridDim.x, blk);
}

template <typename BT, int MODE>
void launch_vec_mode(const VecArgs<BT>& a, unsigned grid, hipStream_t s) {
  const int nv = a.nv;
  if (nv <= 0)
    k_vec<BT, 0, MODE><<<grid, kBlock, 0, s>>>(a);
  else if (nv <= 4)
    k_vec<BT, 4, MODE><<<grid, kBlock, 0, s>>>(a);
  else if (nv <= 8)
    k_vec<BT, 8, MODE><<<grid, kBlock, 0, s>>>(a);
  else if (nv <= 12)
    k_vec<BT, 12, MODE><<<grid, kBlock, 0, s>>>(a);
  else if (nv <= 16)
    k_vec<BT, 16, MODE><<<grid, kBlock, 0, s>>>(a);
  else if (nv <= 20)
    k_vec<BT, 20, MODE><<<grid, kBlock, 0, s>>>(a);
  else if (nv <= 24)
    k_vec<BT, 24, MODE><<<grid, kBlock, 0, s>>>(a);
  else if (nv <= 28)
    k_vec<BT, 28, MODE><<<grid, kBlock, 0, s>>>(a);
  else if (nv <= 32)
    k_vec<BT, 32, MODE><<<grid, kBlock, 0, s>>>(a);
  else
    throw Error(NC_ERR_UNSUPPORTED, "Krylov dimension above 31 is not supported");
}

template <typename BT>
void launch_vec(VecArgs<BT> a, const nc_solver_s* S, hipStream_t s) {
  if (a.N % 4) throw Error(NC_ERR_INTERNAL, "vector passes need a padded length");
  if (!a.xin) throw Error(NC_ERR_INTERNAL, "vector passes need an input vector");
  const unsigned grid = S->red_blocks;
  a.tot = S->fold.p;
  a.counter = reinterpret_cast<unsigned*>(S->fold.p + kPartStrideMax);
  int nvals = 1;
  switch (a.mode) {
    case VM_DOTS: launch_vec_mode<BT, VM_DOTS>(a, grid, s); nvals = a.nv + 1; break;
    case VM_DUAL:
      nvals = 2 * a.nv + 1;
      if (a.nv > 4 && a.nv <= 32 && !a.xout && !a.vout && a.finish != FIN_NONE) {
        if (a.nv <= 8)
          k_vec_dual<BT, 8><<<grid, kBlock, 0, s>>>(a);
        else if (a.nv <= 12)
          k_vec_dual<BT, 12><<<grid, kBlock, 0, s>>>(a);
        else if (a.nv <= 16)
          k_vec_dual<BT, 16><<<grid, kBlock, 0, s>>>(a);
        else if (a.nv <= 20)
          k_vec_dual<BT, 20><<<grid, kBlock, 0, s>>>(a);
        else if (a.nv <= 24)
          k_vec_dual<BT, 24><<<grid, kBlock, 0, s>>>(a);
        else if (a.nv <= 28)
          k_vec_dual<BT, 28><<<grid, kBlock, 0, s>>>(a);
        else
          k_vec_dual<BT, 32><<<grid, kBlock, 0, s>>>(a);
      } else {
        launch_vec_mode<BT, VM_DUAL>(a, grid, s);
      }
      break;
    default:
      if (a.arn_pre) {
        if (a.vout_fp32 != (std::is_same<BT, float>::value ? 1 : 0) || a.xout || !a.vout || a.finish != FIN_NONE)
          throw Error(NC_ERR_INTERNAL, "Arnoldi update pass misconfigured");
        if (a.nv <= 4)
          k_arn_update<BT, 4><<<grid, kBlock, 0, s>>>(a);
        else if (a.nv <= 8)
          k_arn_update<BT, 8><<<grid, kBlock, 0, s>>>(a);
        else if (a.nv <= 12)
          k_arn_update<BT, 12><<<grid, kBlock, 0, s>>>(a);
        else if (a.nv <= 16)
          k_arn_update<BT, 16><<<grid, kBlock, 0, s>>>(a);
        else if (a.nv <= 20)
          k_arn_update<BT, 20><<<grid, kBlock, 0, s>>>(a);
        else if (a.nv <= 24)
          k_arn_update<BT, 24><<<grid, kBlock, 0, s>>>(a);
        else if (a.nv <= 28)
          k_arn_update<BT, 28><<<grid, kBlock, 0, s>>>(a);
        else if (a.nv <= 32)
          k_arn_update<BT, 32><<<grid, kBlock, 0, s>>>(a);
        else
          throw Error(NC_ERR_UNSUPPORTED, "Krylov dimension above 31 is not supported");
      } else {
        launch_vec_mode<BT, VM_UPDATE>(a, grid, s);
      }
      break;
  }
  if (a.finish != FIN_NONE) {
    if (nvals > kPartStrideMax || !a.tot || !a.counter || grid > 4u * kFinBlock)
      throw Error(NC_ERR_INTERNAL, "vector pass fold not set up");
    k_vec_finish<BT><<<nvals, kFinBlock, 0, s>>>(a, grid, nvals);
  }
  NC_CHECK_HIP(hipGetLastError());
}

// ---- Dormand-Prince 5(4) ------------------------------------------------------------
#define NC_DP_A                                                                  \
  {{0, 0, 0, 0, 0, 0},                                                           \
   {1.0 / 5, 0, 0, 0, 0, 0},                                                     \
   {3.0 / 40, 9.0 / 40, 0, 0, 0, 0},                                             \
   {44.0 / 45, -56.0 / 15, 32.0 / 9, 0, 0, 0},                                   \
   {19372.0 / 6561, -25360.0 / 2187, 64448.0 / 6561, -212.0 / 729, 0, 0},        \
   {9017.0 / 3168, -355.0 / 33, 46732.0 / 5247, 49.0 / 176, -5103.0 / 18656, 0}, \
   {35.0 / 384, 0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784, 11.0 / 84}}
__constant__ double c_dp_a[7][6] = NC_DP_A;
const double h_dp_a[7][6] = NC_DP_A;  // the same table for the host (mat-vec epilogues)
#undef NC_DP_A
const double h_dp_c[7] = {0, 1.0 / 5, 3.0 / 10, 4.0 / 5, 8.0 / 9, 1.0, 1.0};
// b - b* (5th minus embedded 4th order weights)
#define NC_DP_E {71.0 / 57600, 0, -71.0 / 16695, 71.0 / 1920, -17253.0 / 339200, 22.0 / 525, -1.0 / 40}
__constant__ double c_dp_e[7] = NC_DP_E;
const double h_dp_e[7] = NC_DP_E;  // the same weights for the host (the mat-vec's error epilogue)
#undef NC_DP_E

// Stage j lives in slot j of k, except that with sw = 1 stages 0 and 6 trade slots (an
// accepted step hands k6 = f(y_{n+1}) on as the next k0 by flipping sw, not by copying).
__device__ __host__ inline int rk_slot(int j, int sw) { return (sw && (j == 0 || j == 6)) ? 6 - j : j; }

// Stage input y_i = y + h * sum_{j<i} a_ij k_j ; writes the mat-vec copy, and at the last
// stage (i = 6, whose input is y_{n+1}) also the FP64 value (ytmp), the only stage input
// the error estimate and the accept read. The stage is a template argument so that its
// STAGE loads of k are issued together (a runtime trip count would wait for each before
// the next); two elements per thread.
template <typename VT, int STAGE>
__global__ __launch_bounds__(kBlock) void k_rk_stage(int64_t N, double h, const double* y, const double* k, int sw,
                                                     double* ytmp, VT* vin) {
  constexpr int R = 2;
  const int64_t stride = (int64_t)gridDim.x * kBlock * R;
  for (int64_t n = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * R; n < N; n += stride) {
    // a_61 = 0: the last stage does not read k1
    Pack<double, R> kv[STAGE];
#pragma unroll
    for (int j = 0; j < STAGE; ++j)
      if (!(STAGE == 6 && j == 1)) kv[j] = *reinterpret_cast<const Pack<double, R>*>(k + (int64_t)rk_slot(j, sw) * N + n);
    const Pack<double, R> yv = *reinterpret_cast<const Pack<double, R>*>(y + n);
    Pack<double, R> o;
    Pack<VT, R> ov;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < STAGE; ++j)
        if (!(STAGE == 6 && j == 1)) acc += c_dp_a[STAGE][j] * kv[j].v[r];
      o.v[r] = yv.v[r] + h * acc;
      ov.v[r] = (VT)o.v[r];
    }
    if constexpr (STAGE == 6) *reinterpret_cast<Pack<double, R>*>(ytmp + n) = o;
    *reinterpret_cast<Pack<VT, R>*>(vin + n) = ov;
  }
}

template <typename VT>
void launch_rk_stage(int stage, unsigned grid, hipStream_t s, int64_t N, double h, const double* y, const double* k,
                     int sw, double* ytmp, VT* vin) {
  switch (stage) {
    case 1: k_rk_stage<VT, 1><<<grid, kBlock, 0, s>>>(N, h, y, k, sw, ytmp, vin); break;
    case 2: k_rk_stage<VT, 2><<<grid, kBlock, 0, s>>>(N, h, y, k, sw, ytmp, vin); break;
    case 3: k_rk_stage<VT, 3><<<grid, kBlock, 0, s>>>(N, h, y, k, sw, ytmp, vin); break;
    case 4: k_rk_stage<VT, 4><<<grid, kBlock, 0, s>>>(N, h, y, k, sw, ytmp, vin); break;
    case 5: k_rk_stage<VT, 5><<<grid, kBlock, 0, s>>>(N, h, y, k, sw, ytmp, vin); break;
    case 6: k_rk_stage<VT, 6><<<grid, kBlock, 0, s>>>(N, h, y, k, sw, ytmp, vin); break;
    default: throw Error(NC_ERR_INTERNAL, "bad RK stage");
  }
}

// Error norm sum_n (h*sum_j e_j k_j / (atol + rtol*max(|y|,|ynew|)))^2 with ynew = ytmp;
// one partial sum per block (folded on the host, which reads them back anyway). With
// R > 0, block 0 also gathers the sink rows ns..ns+R-1 of ynew, k0 and k6 behind the
// partials, so that one copy brings back everything the step controller needs.
__global__ __launch_bounds__(kBlock) void k_rk_err(int64_t N, double h, double atol, double rtol, const double* y,
                                                   const double* ynew, const double* k, int sw, int64_t ns, int R,
                                                   double* partials) {
  if (blockIdx.x == 0 && (int)threadIdx.x < R) {
    double* ex = partials + gridDim.x;
    const int r = threadIdx.x;
    ex[r] = ynew[ns + r];
    ex[R + r] = k[(int64_t)rk_slot(0, sw) * N + ns + r];
    ex[2 * R + r] = k[(int64_t)rk_slot(6, sw) * N + ns + r];
  }
  double acc = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x; n < N; n += stride) {
    double e = 0.0;
    // e_1 = 0 in Dormand-Prince 5(4): k1 is not read
#pragma unroll
    for (int j = 0; j < 7; ++j)
      if (j != 1) e += c_dp_e[j] * k[(int64_t)rk_slot(j, sw) * N + n];
    e *= h;
    const double sc = atol + rtol * fmax(fabs(y[n]), fabs(ynew[n]));
    const double r = e / sc;
    acc += r * r;
  }
  __shared__ double red[kBlock / 64];
  double v = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) s += red[w];
    partials[blockIdx.x] = s;
  }
}

// Cubic Hermite dense output on an accepted step, written over y:
// y(t + theta h) from y, y1 = y(t + h) and the derivatives k0 = f(y), k6 = f(y1).
__global__ __launch_bounds__(kBlock) void k_rk_hermite(int64_t N, double theta, double h, double* y, const double* y1,
                                                       const double* k0, const double* k6) {
  const double t2 = theta * theta, t3 = t2 * theta;
  const double a0 = 2 * t3 - 3 * t2 + 1, b0 = (t3 - 2 * t2 + theta) * h;
  const double a1 = -2 * t3 + 3 * t2, b1 = (t3 - t2) * h;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x; n < N; n += stride)
    y[n] = a0 * y[n] + b0 * k0[n] + a1 * y1[n] + b1 * k6[n];
}

// Copy the sink rows (indices n_states .. n_states+R-1) of the m+1 basis vectors and of
// the state u into the kbuf sink region, so the host can evaluate the sink-mass budget of
// any combination of the basis.
template <typename BT>
__global__ void k_gather_sinks(const BT* V, int64_t ld, int nvec, const double* u, int64_t n_states, int R,
                               double* dst) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (nvec + 1) * R) return;
  const int j = t / R, r = t % R;
  dst[t] = j < nvec ? (double)V[(int64_t)j * ld + n_states + r] : u[n_states + r];
}

__global__ void k_copy_d(const double* in, double* out, int64_t N) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x; n < N; n += stride) out[n] = in[n];
}

template <typename VT>
__global__ void k_to_vt(const double* in, VT* out, int64_t N) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x; n < N; n += stride) out[n] = (VT)in[n];
}

double round_step(double x) {
  // Expokit rounds step sizes to 2 significant digits
  if (!(x > 0.0)) return x;
  const double s = std::pow(10.0, std::floor(std::log10(x)) - 1.0);
  return std::ceil(x / s) * s;
}

}  // namespace

void solver_reduce_setup(nc_solver_s* S, int nvals) {
  (void)nvals;
  // Grid of the vector passes: 1024 blocks of 256 threads for a Krylov basis stored in
  // FP32 and 512 otherwise (on the 1024^2 toggle-switch solve, 256 to 4096 blocks are
  // within 1% of each other since the passes issue their basis loads in batches)
  const unsigned vb = (S->method == NC_METHOD_KRYLOV && S->vs == 4) ? 1024u : kVecBlocks;
  S->red_blocks = std::min<unsigned>(vb, grid_for(S->Np / 4, kBlock));
  // Krylov on one block: the projection sums go over 512-row blocks of the states (one
  // workgroup each, the mat-vec's own row workgroups), folded with the sink rows
  S->dblk = 0;
  if (S->method == NC_METHOD_KRYLOV && !S->SA && S->blocks == 1) {
    const int64_t nb = (S->A->n + kDualRows - 1) / kDualRows;
    if (nb >= 1 && nb <= kDualMaxBlocks) S->dblk = nb;
  }
  // the RK error pass puts 3 R gathered sink values behind its partials
  S->partials.reserve(std::max<size_t>({(size_t)S->red_blocks * kPartStrideMax, (size_t)S->dblk * kPartStrideMax,
                                        (size_t)S->red_blocks + 3 * kMaxReactions}));
  if (!S->fold.p) {
    // folded totals, then the fold's arrival counter (zero between passes)
    S->fold.alloc(kPartStrideMax + 1);
    NC_CHECK_HIP(hipMemsetAsync(S->fold.p, 0, sizeof(double) * (kPartStrideMax + 1), S->s));
  }
  if (!S->hpart) {
    // partials (up to kMaxRedBlocks, or one per row workgroup and sink row of the mat-vec's
    // error epilogue) and the RK step's 3 R gathered sink values, then the carried sinks of
    // u and the step's three sink sets (4 R) behind them
    S->hpart_cap = std::max<int64_t>(kMaxRedBlocks, matvec_err_parts(S->A));
    S->hpart = static_cast<double*>(pinned_alloc(sizeof(double) * (S->hpart_cap + 7 * kMaxReactions)));
    NC_CHECK_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&S->hpart_dev), S->hpart, 0));
  }
}

// The right-hand-side operator: the FSP generator, or the block sensitivity operator.
static void op_apply(nc_solver_s* S, double t, const void* v, double* out) {
  if (S->SA)
    sens_matvec_launch(S->SA, t, v, S->Npb, out, S->Npb, S->s);
  else
    matvec_launch(S->A, t, v, out, false, S->s, true);
}

// Device (padded, block stride Npb) <-> host (contiguous, block stride Nb) copies.
static void copy_from_host(nc_solver_s* S, double* dev, const double* host) {
  if (S->blocks == 1)
    NC_CHECK_HIP(hipMemcpyAsync(dev, host, sizeof(double) * S->N, hipMemcpyHostToDevice, S->s));
  else
    NC_CHECK_HIP(hipMemcpy2DAsync(dev, sizeof(double) * S->Npb, host, sizeof(double) * S->Nb, sizeof(double) * S->Nb,
                                  S->blocks, hipMemcpyHostToDevice, S->s));
}
static void copy_to_host(nc_solver_s* S, double* host, const double* dev) {
  if (S->blocks == 1)
    NC_CHECK_HIP(hipMemcpyAsync(host, dev, sizeof(double) * S->N, hipMemcpyDeviceToHost, S->s));
  else
    NC_CHECK_HIP(hipMemcpy2DAsync(host, sizeof(double) * S->Nb, dev, sizeof(double) * S->Npb, sizeof(double) * S->Nb,
                                  S->blocks, hipMemcpyDeviceToHost, S->s));
  NC_CHECK_HIP(hipStreamSynchronize(S->s));
}

// ---- Krylov expv ---------------------------------------------------------------------
// Arnoldi column j in two passes over the stored basis V_0..V_j (FP32 or FP64, kept
// unnormalised with scales sc_i):
//   pass 1 (VM_DUAL):   V_i.q and V_i.V_j for q = A V_j, folded into per-value totals
//   pass 2 (VM_UPDATE): every workgroup first sets |V_j|, the Gram column of V_j and solves
//                       G h = a, a_i = sc_i sc_j V_i.q, from the totals (arnoldi_prologue),
//                       then V_{j+1} = sc_j q - sum_i h_i sc_i V_i
// Projecting with the Gram matrix of the stored vectors makes V_{j+1} orthogonal to them
// to FP64 rounding in one projection: the work a second classical Gram-Schmidt pass
// would do, without a third pass over V.
// q = A V_{nv-1} and the projection sums of V_0..V_{nv-1} (vbase) over 512-row blocks:
// by the mat-vec's projection epilogue, or by the block pass over the stored q (the same
// arithmetic), then folded with the sink rows by `finish`.
template <typename BT>
static void dual_blocks_pass(nc_solver_s* S, const BT* vbase, int nv, int finish) {
  nc_matrix_s* A = S->A;
  const BT* vlast = vbase + (int64_t)(nv - 1) * S->Np;
  if (matvec_dots_fusable(A)) {
    const DualDots d{vbase, S->Np, nv, nullptr, S->partials.p, S->dblk};
    matvec_launch_dots(A, vlast, S->q.p, S->s, d);
  } else {
    op_apply(S, 0.0, vlast, S->q.p);
    k_dual_blocks<BT><<<(unsigned)S->dblk, kBlock, 0, S->s>>>(vbase, S->Np, nv, S->q.p, A->n, S->partials.p);
    NC_CHECK_HIP(hipGetLastError());
  }
  VecArgs<BT> a{};
  a.N = S->Np;
  a.nv = nv;
  a.V = vbase;
  a.xin = S->q.p;
  a.partials = S->partials.p;
  a.kb = S->kbuf.p;
  a.L = KLayout{S->m};
  a.btol = kBreakdownTol;
  a.finish = finish;
  a.tot = S->fold.p;
  a.counter = reinterpret_cast<unsigned*>(S->fold.p + kPartStrideMax);
  a.tail0 = A->n;
  a.tail1 = A->n + A->nr;
  const int nvals = 2 * nv + 1;
  if (nvals > kPartStrideMax || S->dblk > 4 * kFinBlock || a.tail1 > S->Np)
    throw Error(NC_ERR_INTERNAL, "projection blocks not set up");
  k_vec_finish<BT><<<nvals, kFinBlock, 0, S->s>>>(a, (unsigned)S->dblk, nvals);
  NC_CHECK_HIP(hipGetLastError());
}

template <typename BT>
static void arnoldi_iteration(nc_solver_s* S, int j) {
  const int64_t N = S->Np;
  const KLayout L{S->m};
  hipStream_t s = S->s;
  auto Vp = [&](int i) { return reinterpret_cast<BT*>(S->V.p + (size_t)i * N * S->vs); };
  if (S->dblk) {
    dual_blocks_pass<BT>(S, Vp(0), j + 1, FIN_FOLD);
    VecArgs<BT> a{};
    a.N = N;
    a.nv = j + 1;
    a.V = Vp(0);
    a.partials = S->partials.p;
    a.kb = S->kbuf.p;
    a.gram = S->gram.p;
    a.L = L;
    a.j = j;
    a.btol = kBreakdownTol;
    a.vout_fp32 = std::is_same<BT, float>::value ? 1 : 0;
    a.xin = S->q.p;
    a.mode = VM_UPDATE;
    a.arn_pre = 1;
    a.vout = Vp(j + 1);
    a.finish = FIN_NONE;
    launch_vec(a, S, s);
    return;
  }
  op_apply(S, 0.0, Vp(j), S->q.p);
  VecArgs<BT> a{};
  a.N = N;
  a.nv = j + 1;
  a.V = Vp(0);
  a.partials = S->partials.p;
  a.kb = S->kbuf.p;
  a.gram = S->gram.p;
  a.L = L;
  a.j = j;
  a.btol = kBreakdownTol;
  a.vout_fp32 = std::is_same<BT, float>::value ? 1 : 0;
  a.mode = VM_DUAL;
  a.xin = S->q.p;
  a.finish = FIN_FOLD;
  launch_vec(a, S, s);
  // the update applies the finishing step (|V_j|, Gram column, h) from the folded totals
  a.mode = VM_UPDATE;
  a.arn_pre = 1;
  a.vout = Vp(j + 1);
  a.finish = FIN_NONE;
  launch_vec(a, S, s);
}

template <typename BT>
static void arnoldi_last(nc_solver_s* S) {
  // |V_m| (H(m, m-1), sc_m) and avnorm = |A v_m| in one VM_DUAL pass over V_m
  const KLayout L{S->m};
  BT* vm = reinterpret_cast<BT*>(S->V.p + (size_t)S->m * S->Np * S->vs);
  if (S->dblk) {
    dual_blocks_pass<BT>(S, vm, 1, FIN_LAST);
    return;
  }
  op_apply(S, 0.0, vm, S->q.p);
  VecArgs<BT> a{};
  a.N = S->Np;
  a.nv = 1;
  a.V = vm;
  a.mode = VM_DUAL;
  a.xin = S->q.p;
  a.partials = S->partials.p;
  a.kb = S->kbuf.p;
  a.L = L;
  a.btol = kBreakdownTol;
  a.finish = FIN_LAST;
  launch_vec(a, S, S->s);
}

// The device work of one Arnoldi process (m + 1 mat-vecs, the projections and the
// |A v_m| norm). Its launches and arguments are identical for every Krylov step, so it is
// captured once into a HIP graph and replayed.
static void arnoldi_launch(nc_solver_s* S) {
  const KLayout L{S->m};
  const bool fp32 = S->A->precision == NC_FP32;
  // V_0 (raw) and sc[0] = 1/beta are already on the device
  for (int j = 0; j < S->m; ++j) {
    if (fp32)
      arnoldi_iteration<float>(S, j);
    else
      arnoldi_iteration<double>(S, j);
  }
  if (fp32)
    arnoldi_last<float>(S);
  else
    arnoldi_last<double>(S);
  const int R = S->A->nr;
  double* dst = S->kbuf.p + L.sinkv(0, 0, R);
  if (fp32)
    k_gather_sinks<float><<<grid_for((int64_t)(S->m + 2) * R, kBlock), kBlock, 0, S->s>>>(reinterpret_cast<const float*>(S->V.p), S->Np, S->m + 1,
                                                       S->u.p, S->A->n, R, dst);
  else
    k_gather_sinks<double><<<grid_for((int64_t)(S->m + 2) * R, kBlock), kBlock, 0, S->s>>>(reinterpret_cast<const double*>(S->V.p), S->Np, S->m + 1,
                                                        S->u.p, S->A->n, R, dst);
  NC_CHECK_HIP(hipGetLastError());
}

static void arnoldi_run(nc_solver_s* S) {
  const bool use_graph = !S->A->timing && S->opts.no_graphs == 0;
  if (!use_graph) {
    arnoldi_launch(S);
    return;
  }
  if (S->arnoldi_exec && S->arnoldi_gen != S->A->launch_gen) {
    // the matrix's launch shape changed since the capture (autotune, set_rows_per_thread)
    NC_CHECK_HIP(hipStreamSynchronize(S->s));
    (void)hipGraphExecDestroy(S->arnoldi_exec);
    (void)hipGraphDestroy(S->arnoldi_graph);
    S->arnoldi_exec = nullptr;
    S->arnoldi_graph = nullptr;
  }
  if (!S->arnoldi_exec) {
    hipGraph_t g = nullptr;
    NC_CHECK_HIP(hipStreamBeginCapture(S->s, hipStreamCaptureModeThreadLocal));
    try {
      arnoldi_launch(S);
    } catch (...) {
      (void)hipStreamEndCapture(S->s, &g);
      if (g) (void)hipGraphDestroy(g);
      (void)hipGetLastError();
      throw;
    }
    NC_CHECK_HIP(hipStreamEndCapture(S->s, &g));
    NC_CHECK_HIP(hipGraphInstantiate(&S->arnoldi_exec, g, nullptr, nullptr, 0));
    S->arnoldi_graph = g;
    S->arnoldi_gen = S->A->launch_gen;
  }
  NC_CHECK_HIP(hipGraphLaunch(S->arnoldi_exec, S->s));
}

// Kernel timing of one solver's advance. The generator may be shared by several solvers
// (the flag and the event pool live on the matrix), so the solver's own time_kernels
// setting applies only while it advances, and the timings recorded meanwhile are
// credited to its own statistics.
struct TimingScope {
  nc_solver_s* S;
  bool was, was_plain;
  double ms0;
  int64_t n0;
  explicit TimingScope(nc_solver_s* s_) : S(s_), was(s_->A->timing), was_plain(s_->A->timing_plain) {
    S->A->timing = S->opts.time_kernels != 0;
    S->A->timing_plain = S->opts.time_kernels == 2;
    ms0 = S->A->timed_ms;
    n0 = S->A->timed_count;
  }
  ~TimingScope() {
    try {
      matvec_harvest_timing(S->A);
    } catch (...) {
    }
    S->st.matvec_ms += S->A->timed_ms - ms0;
    S->st.matvecs_timed += S->A->timed_count - n0;
    S->A->timing = was;
    S->A->timing_plain = was_plain;
  }
};

// Sink-mass budget of the adaptive solve (fspsolve.jl:145-148): g(t) = sum(sinks) - c t.
static double budget(const EventSpec* ev, double sink_sum, double t) {
  return sink_sum - ev->fsptol * t / ev->tend;
}

void krylov_advance(nc_solver_s* S, double t_target, EventSpec* ev, bool one_step) {
  TimingScope timing_scope(S);
  nc_matrix_s* A = S->A;
  const int m = S->m;
  const int64_t N = S->Np;
  const KLayout L{m};
  const int R = A->nr;
  hipStream_t s = S->s;
  const bool fp32 = A->precision == NC_FP32;
  const double tol = S->opts.krylov_tol;
  const double btol = kBreakdownTol, gamma = 0.9, delta = 1.2;
  const int mxrej = 10;
  auto Vp = [&](int i) -> unsigned char* { return S->V.p + (size_t)i * N * S->vs; };
  if (!S->started) {
    // first step size from the norm estimate (Expokit)
    S->anorm = std::max(S->SA ? S->SA->norm1 : A->norm1, 1e-300);
    const double fact = std::pow((m + 1) / std::exp(1.0), m + 1) * std::sqrt(2.0 * M_PI * (m + 1));
    const double b = std::max(S->beta, 1e-300);
    S->t_new = round_step((1.0 / S->anorm) * std::pow((fact * tol) / (4.0 * b * S->anorm), 1.0 / m));
    S->started = true;
  }
  const double t_eps = 1e-13 * std::max(1.0, std::fabs(t_target));
  while (S->t < t_target - t_eps) {
    if (S->opts.max_steps && S->st.steps >= S->opts.max_steps)
      throw Error(NC_ERR_NUMERIC, "Krylov integrator exceeded max_steps");
    if (S->beta == 0.0 && !S->beta_pending) {  // zero vector stays zero
      S->t = t_target;
      break;
    }
    double t_step = std::min(t_target - S->t, S->t_new);
    arnoldi_run(S);
    S->st.matvecs += m + 1;
    NC_CHECK_HIP(hipMemcpyAsync(S->hbuf, S->kbuf.p, sizeof(double) * L.size_with_sinks(R), hipMemcpyDeviceToHost, s));
    NC_CHECK_HIP(hipStreamSynchronize(s));
    S->st.host_syncs++;
    matvec_harvest_timing(A);
    const double* hb = S->hbuf;
    if (S->beta_pending) {
      S->beta = hb[L.misc(MISC_BETA)];
      S->beta_pending = false;
    }
    // happy breakdown detection
    int mb = m, k1 = 2;
    for (int j = 0; j < m; ++j) {
      if (hb[L.H(j + 1, j)] < btol) {
        mb = j + 1;
        k1 = 0;
        t_step = t_target - S->t;
        break;
      }
    }
    const double avnorm = hb[L.misc(MISC_AVN)];
    const int mxf = mb + k1;  // size of the exponentiated matrix
    std::vector<double> Hm((size_t)mxf * mxf, 0.0), F(mxf);  // F: first column of exp(tau Hm)
    auto exp_tau = [&](double tau, std::vector<double>& out) {
      for (int i = 0; i < mxf; ++i)
        for (int jj = 0; jj < mxf; ++jj) {
          double v = 0.0;
          if (i < mb + 1 && jj < mb && i <= jj + 1 && i < m + 1) v = hb[L.H(i, jj)];
          if (k1 != 0 && i == m + 1 && jj == m) v = 1.0;
          Hm[(size_t)i * mxf + jj] = tau * v;
        }
      expm_first_column(mxf, Hm.data(), out.data());
    };
    double err_loc = btol, xm = 1.0 / m;
    for (int ireject = 0;; ++ireject) {
      exp_tau(t_step, F);
      if (k1 == 0) {
        err_loc = btol;
        break;
      }
      const double phi1 = std::fabs(S->beta * F[m]);
      const double phi2 = std::fabs(S->beta * F[m + 1] * avnorm);
      if (phi1 > 10.0 * phi2) {
        err_loc = phi2;
        xm = 1.0 / m;
      } else if (phi1 > phi2) {
        err_loc = (phi1 * phi2) / (phi1 - phi2);
        xm = 1.0 / m;
      } else {
        err_loc = phi1;
        xm = 1.0 / (m - 1);
      }
      if (err_loc <= delta * t_step * tol) break;
      if (ireject == mxrej) throw Error(NC_ERR_NUMERIC, "Krylov step rejected too many times; tolerance too tight");
      t_step = round_step(gamma * t_step * std::pow(t_step * tol / err_loc, xm));
      S->st.rejected++;
    }
    const int mx = mb + std::max(0, k1 - 1);
    // sink mass of w(tau) = beta V exp(tau H) e1, with the exact state for the first term
    auto sink_sum = [&](const std::vector<double>& Fm) {
      double tot = 0.0;
      for (int r = 0; r < R; ++r) {
        double v = Fm[0] * hb[L.sinkv(m + 1, r, R)];
        for (int i = 1; i < mx; ++i) v += S->beta * Fm[i] * hb[L.sc(i)] * hb[L.sinkv(i, r, R)];
        tot += v;
      }
      return tot;
    };
    bool fire = false;
    double sinks0 = 0.0;  // sink mass of the state at the start of the step
    for (int r = 0; r < R; ++r) sinks0 += hb[L.sinkv(m + 1, r, R)];
    if (ev && R > 0 && budget(ev, sinks0, S->t) <= 0.0 && budget(ev, sink_sum(F), S->t + t_step) > 0.0) {
      // the budget is crossed inside this step: locate the crossing on the exact Krylov
      // solution w(tau) by bisection (the reference's callback root-finds to eps)
      std::vector<double> Ft(mxf);
      double lo = 0.0, hi = t_step;
      for (int it = 0; it < 200 && hi - lo > 4.0 * std::numeric_limits<double>::epsilon() * std::max(1.0, S->t + hi); ++it) {
        const double mid = 0.5 * (lo + hi);
        exp_tau(mid, Ft);
        if (budget(ev, sink_sum(Ft), S->t + mid) > 0.0)
          hi = mid;
        else
          lo = mid;
      }
      t_step = hi;
      exp_tau(t_step, F);
      fire = true;
    }
    // w_new = c0/beta * w + sum_{i>=1} c_i sc_i V_i ; c = beta * F[:,0]
    double* coef = S->hcoef;  // pinned; the previous upload completed at the last sync
    std::fill(coef, coef + m + 2, 0.0);
    for (int i = 1; i < mx; ++i) coef[i - 1] = S->beta * F[i];
    coef[m + 1] = F[0];  // (beta*F00)/beta
    NC_CHECK_HIP(hipMemcpyAsync(S->coef.p, coef, sizeof(double) * (m + 2), hipMemcpyHostToDevice, s));
    S->t += t_step;
    const bool hit = !fire && S->t >= t_target - t_eps;
    if (hit) S->t = t_target;
    {
      auto fill = [&](auto& a, auto* vbase) {
        a.N = N;
        a.nv = mx - 1;
        a.V = vbase;
        a.sc = S->kbuf.p + L.sc(1);
        a.coef = S->coef.p;
        a.csign = 1.0;
        a.xin = S->u.p;
        a.xscale = S->coef.p + (m + 1);
        a.xout = S->u.p;
        a.vout = Vp(0);
        a.vout_fp32 = fp32 ? 1 : 0;
        a.partials = S->partials.p;
        a.kb = S->kbuf.p;
        a.L = L;
        a.finish = FIN_BETA;
        a.mode = VM_UPDATE;
      };
      if (fp32) {
        VecArgs<float> a{};
        fill(a, reinterpret_cast<const float*>(Vp(1)));
        launch_vec(a, S, s);
      } else {
        VecArgs<double> a{};
        fill(a, reinterpret_cast<const double*>(Vp(1)));
        launch_vec(a, S, s);
      }
    }
    // the new beta (written on the device by the combination) arrives with the next H
    // download, so no synchronisation is needed here
    S->beta_pending = true;
    S->st.steps++;
    err_loc = std::max(err_loc, 1e-300);
    S->st.error_estimate += std::max(err_loc, S->anorm * 2.220446049250313e-16);
    if (k1 != 0 && !fire) S->t_new = round_step(gamma * t_step * std::pow(t_step * tol / err_loc, xm));
    if (fire) {
      ev->fired = true;
      break;
    }
    if (hit || one_step) break;
  }
}

// ---- Dormand-Prince 5(4) ---------------------------------------------------------------
void rk_advance(nc_solver_s* S, double t_target, EventSpec* ev, bool one_step) {
  TimingScope timing_scope(S);
  nc_matrix_s* A = S->A;
  const int64_t N = S->Np;  // padded: rows beyond the true length stay zero
  const int R = A->nr;
  const int64_t ns = A->n;
  hipStream_t s = S->s;
  const bool fp32 = A->precision == NC_FP32;
  const double rtol = S->opts.rtol, atol = S->opts.atol;
  const unsigned g = std::min<unsigned>(4096, grid_for(N, kBlock));  // elementwise stage kernels
  const unsigned g2 = std::min<unsigned>(4096, grid_for(N / 2, kBlock));  // two elements per thread
  double* k = S->k.p;
  auto kp = [&](int j) { return k + (int64_t)rk_slot(j, S->ksw) * N; };
  auto to_vin = [&](const double* src) {
    if (fp32)
      k_to_vt<float><<<g, kBlock, 0, s>>>(src, reinterpret_cast<float*>(S->rkv.p), N);
    else
      k_to_vt<double><<<g, kBlock, 0, s>>>(src, reinterpret_cast<double*>(S->rkv.p), N);
  };
  // The mat-vec that forms k_s also forms stage s + 1's input in its epilogue (the stored
  // generator at 2 rows per thread, and the sensitivity operator up to 16 blocks; the stage
  // pass runs on its own for stage 1 and for the computed modes);
  // the stage inputs alternate between two buffers, since a mat-vec gathers from its input
  // while its epilogue writes the next one.
  const bool fuse = S->rkv2.p && (S->SA ? sens_rk_fusable(S->SA) : matvec_rk_fusable(A));
  // the error pass in the epilogue of the step's last mat-vec (plain generator)
  const int64_t err_parts = matvec_err_parts(A);
  const bool efuse = fuse && !S->SA && err_parts <= S->hpart_cap;
  const int Rg = (ev && R > 0) ? R : 0;
  auto stages = [&](double h) {
    unsigned char* cur = S->rkv.p;
    unsigned char* nxt = S->rkv2.p;
    for (int stage = 1; stage < 7; ++stage) {
      if (stage == 1 || !fuse) {
        if (fp32)
          launch_rk_stage<float>(stage, g2, s, N, h, S->u.p, k, S->ksw, S->ytmp.p, reinterpret_cast<float*>(cur));
        else
          launch_rk_stage<double>(stage, g2, s, N, h, S->u.p, k, S->ksw, S->ytmp.p, reinterpret_cast<double*>(cur));
        NC_CHECK_HIP(hipGetLastError());
      }
      const double ts = S->t + h_dp_c[stage] * h;
      if (fuse && stage < 6) {
        const int next = stage + 1;
        RkEpilogue e{};
        for (int j = 0; j < stage; ++j) {
          if (next == 6 && j == 1) continue;  // a_61 = 0: the last stage does not read k1
          e.mask |= 1 << j;
          e.k[j] = kp(j);
          e.a[j] = h_dp_a[next][j];
        }
        e.k[0] = kp(0);
        e.a_new = h_dp_a[next][stage];
        e.h = h;
        e.u = S->u.p;
        e.ytmp = next == 6 ? S->ytmp.p : nullptr;
        e.vin = nxt;
        if (S->SA)
          sens_matvec_launch(S->SA, ts, cur, S->Npb, kp(stage), S->Npb, s, &e);
        else
          matvec_launch_rk(A, ts, cur, kp(stage), s, e);
        std::swap(cur, nxt);
      } else if (efuse && stage == 6) {
        RkErrEpilogue e{};
        for (int j = 0; j < 6; ++j) e.k[j] = kp(j);
        for (int j = 0; j < 7; ++j) e.e[j] = h_dp_e[j];
        e.h = h;
        e.atol = atol;
        e.rtol = rtol;
        e.u = S->u.p;
        e.ynew = S->ytmp.p;
        e.part = S->hpart_dev;
        e.ex = S->hpart_dev + err_parts;
        e.n_ex = Rg;
        matvec_launch_err(A, ts, cur, kp(stage), s, e);
      } else {
        op_apply(S, ts, cur, kp(stage));
      }
      S->st.matvecs++;
    }
  };
  if (!S->fsal_valid) {
    to_vin(S->u.p);
    op_apply(S, S->t, S->rkv.p, kp(0));
    S->st.matvecs++;
    S->fsal_valid = true;
  }
  if (S->h <= 0.0) {
    // initial step (Hairer, Norsett & Wanner, Solving ODEs I, II.4): based on |f(t0,y0)|
    std::vector<double> sums(2);
    VecArgs<double> a{};
    a.N = N;
    a.nv = 0;
    a.xin = kp(0);
    a.partials = S->partials.p;
    a.kb = S->kbuf.p;
    a.L = KLayout{S->m};
    a.finish = FIN_SUMS;
    launch_vec(a, S, s);
    NC_CHECK_HIP(hipMemcpyAsync(sums.data(), S->kbuf.p, sizeof(double), hipMemcpyDeviceToHost, s));
    NC_CHECK_HIP(hipStreamSynchronize(s));
    const double fnorm = std::sqrt(sums[0]);
    S->h = S->opts.h0 > 0 ? S->opts.h0 : (fnorm > 0 ? 0.01 / fnorm : 1e-3);
    S->h = std::max(S->h, 1e-12);
  }
  // pinned: sinks of u, y_{n+1}, k0, k6 (R each), behind the readback region
  double* hs = S->hpart + S->hpart_cap + 3 * kMaxReactions;
  if (ev && R > 0) {
    NC_CHECK_HIP(hipMemcpyAsync(hs, S->u.p + ns, sizeof(double) * R, hipMemcpyDeviceToHost, s));
    NC_CHECK_HIP(hipStreamSynchronize(s));
  }
  const double t_eps = 1e-13 * std::max(1.0, std::fabs(t_target));
  double errsum = 0.0;
  while (S->t < t_target - t_eps) {
    if (S->opts.max_steps && S->st.steps + S->st.rejected >= S->opts.max_steps)
      throw Error(NC_ERR_NUMERIC, "RK integrator exceeded max_steps");
    double h = std::min(S->h, t_target - S->t);
    const bool clamped = h < S->h;
    stages(h);
    // after stage 6 the stage input is y + h*sum b_j k_j (row 7 of A): ytmp = y_{n+1}
    // the partials and the gathered sinks go straight to the pinned host block: the step's
    // read-back needs no copy launch, only the synchronisation
    if (!efuse) {
      k_rk_err<<<S->red_blocks, kBlock, 0, s>>>(N, h, atol, rtol, S->u.p, S->ytmp.p, k, S->ksw, ns, Rg, S->hpart_dev);
      NC_CHECK_HIP(hipGetLastError());
    }
    NC_CHECK_HIP(hipStreamSynchronize(s));
    const int64_t np = efuse ? err_parts : (int64_t)S->red_blocks;
    for (int r = 0; r < Rg; ++r) {  // sinks of y_{n+1}, k0 and k6 behind the partials
      hs[R + r] = S->hpart[np + r];
      hs[2 * R + r] = S->hpart[np + R + r];
      hs[3 * R + r] = S->hpart[np + 2 * R + r];
    }
    // four running sums (the epilogue's partials number one per row workgroup), then added
    double es[4] = {0.0, 0.0, 0.0, 0.0};
    int64_t b = 0;
    for (; b + 4 <= np; b += 4)
      for (int q = 0; q < 4; ++q) es[q] += S->hpart[b + q];
    for (; b < np; ++b) es[0] += S->hpart[b];
    errsum = (es[0] + es[1]) + (es[2] + es[3]);
    S->st.host_syncs++;
    matvec_harvest_timing(A);
    const double err = std::sqrt(errsum / (double)S->N);
    if (!std::isfinite(err)) throw Error(NC_ERR_NUMERIC, "non-finite RK error estimate");
    double fac = err > 0 ? 0.9 * std::pow(err, -0.2) : 5.0;
    if (err <= 1.0) {
      if (ev && R > 0) {
        double s0 = 0, s1 = 0, d0 = 0, d1 = 0;
        for (int r = 0; r < R; ++r) s0 += hs[r], s1 += hs[R + r], d0 += hs[2 * R + r], d1 += hs[3 * R + r];
        const double c = ev->fsptol / ev->tend;
        const double g0 = s0 - c * S->t, g1 = s1 - c * (S->t + h);
        if (g0 <= 0.0 && g1 > 0.0) {
          // cubic Hermite dense output of g on the step (derivatives from the FSAL stages),
          // crossing located by bisection; the event state is the same Hermite
          // interpolant of the whole state, so its sink mass is the root-found one
          const double dg0 = (d0 - c) * h, dg1 = (d1 - c) * h;
          auto herm = [&](double th) {
            const double t2 = th * th, t3 = t2 * th;
            return (2 * t3 - 3 * t2 + 1) * g0 + (t3 - 2 * t2 + th) * dg0 + (-2 * t3 + 3 * t2) * g1 + (t3 - t2) * dg1;
          };
          double lo = 0.0, hi = 1.0;
          for (int it = 0; it < 200 && (hi - lo) * h > 4.0 * std::numeric_limits<double>::epsilon() * std::max(1.0, S->t + h); ++it) {
            const double mid = 0.5 * (lo + hi);
            if (herm(mid) > 0.0)
              hi = mid;
            else
              lo = mid;
          }
          k_rk_hermite<<<g, kBlock, 0, s>>>(N, hi, h, S->u.p, S->ytmp.p, kp(0), kp(6));
          NC_CHECK_HIP(hipGetLastError());
          S->t += hi * h;
          S->st.steps++;
          S->fsal_valid = false;  // k0 no longer matches u
          ev->fired = true;
          break;
        }
      }
      // accept: u <- ytmp, k0 <- k6 (FSAL), by exchanging buffers
      std::swap(S->u, S->ytmp);
      S->ksw ^= 1;
      S->t += h;
      S->st.steps++;
      if (S->t > t_target - t_eps) S->t = t_target;
      for (int r = 0; ev && r < R; ++r) hs[r] = hs[R + r];
      fac = std::min(5.0, std::max(0.2, fac));
      if (!clamped || fac < 1.0) S->h = h * fac;
      if (one_step) break;
    } else {
      S->st.rejected++;
      S->h = h * std::max(0.2, fac);
    }
    if (S->h < 1e-14 * std::max(1.0, std::fabs(S->t))) throw Error(NC_ERR_NUMERIC, "RK step size underflow");
  }
}

}  // namespace nc

using namespace nc;

namespace {

static nc_solver_s* solver_create_impl(nc_matrix_s* A, const nc_solve_opts* opts, nc_sensmatrix_s* SA = nullptr) {
  {
    NC_REQUIRE(A, NC_ERR_ARG, "null matrix");
    auto S = new nc_solver_s();
    try {
      if (opts)
        S->opts = *opts;
      else
        nc_solve_opts_default(&S->opts);
      if (S->opts.krylov_m <= 0) S->opts.krylov_m = 30;
      if (S->opts.krylov_tol <= 0) S->opts.krylov_tol = 1e-12;
      if (S->opts.rtol <= 0) S->opts.rtol = 1e-8;
      if (S->opts.atol <= 0) S->opts.atol = 1e-14;
      NC_REQUIRE(S->opts.krylov_m >= 2 && S->opts.krylov_m <= 31, NC_ERR_ARG, "krylov_m must be in 2..31");
      S->A = A;
      S->s = A->stream.s;
      S->SA = SA;
      S->blocks = SA ? SA->P + 1 : 1;
      S->Nb = A->n_rows;
      S->Npb = (S->Nb + 1023) / 1024 * 1024;
      S->N = S->blocks * S->Nb;
      S->vs = A->vsize;
      const bool tv = !A->tv_r.empty();
      int method = S->opts.method;
      if (method == NC_METHOD_AUTO) method = tv ? NC_METHOD_RK45 : NC_METHOD_KRYLOV;
      NC_REQUIRE(!(method == NC_METHOD_KRYLOV && tv), NC_ERR_ARG,
                 "the Krylov integrator needs a time-invariant generator; use NC_METHOD_RK45");
      S->method = method;
      S->m = S->opts.krylov_m;
      S->Np = S->blocks * S->Npb;
      S->u.alloc(S->Np);
      NC_CHECK_HIP(hipMemsetAsync(S->u.p, 0, sizeof(double) * S->Np, S->s));
      const KLayout L{S->m};
      const int ksize = L.size_with_sinks(A->nr);
      S->kbuf.alloc(ksize);
      NC_CHECK_HIP(hipMemsetAsync(S->kbuf.p, 0, sizeof(double) * ksize, S->s));
      S->hbuf = static_cast<double*>(pinned_alloc(sizeof(double) * ksize));
      S->hcoef = static_cast<double*>(pinned_alloc(sizeof(double) * (S->m + 2)));
      // k_vec writes partials with a stride of MAXV+1 <= 33 doubles per block
      solver_reduce_setup(S, 32);
      if (method == NC_METHOD_KRYLOV) {
        S->V.alloc((size_t)(S->m + 1) * S->Np * S->vs);
        S->q.alloc(S->Np);
        NC_CHECK_HIP(hipMemsetAsync(S->V.p, 0, (size_t)(S->m + 1) * S->Np * S->vs, S->s));
        NC_CHECK_HIP(hipMemsetAsync(S->q.p, 0, sizeof(double) * S->Np, S->s));
        S->coef.alloc(S->m + 2);
        S->gram.alloc((size_t)(S->m + 1) * (S->m + 1));
        NC_CHECK_HIP(hipMemsetAsync(S->gram.p, 0, sizeof(double) * (S->m + 1) * (S->m + 1), S->s));
      } else {
        S->rkv.alloc(S->Np * S->vs);
        S->k.alloc((size_t)7 * S->Np);
        S->ytmp.alloc(S->Np);
        NC_CHECK_HIP(hipMemsetAsync(S->rkv.p, 0, S->Np * S->vs, S->s));
        S->rkv2.alloc(S->Np * S->vs);
        NC_CHECK_HIP(hipMemsetAsync(S->rkv2.p, 0, S->Np * S->vs, S->s));
        NC_CHECK_HIP(hipMemsetAsync(S->k.p, 0, sizeof(double) * 7 * S->Np, S->s));
        NC_CHECK_HIP(hipMemsetAsync(S->ytmp.p, 0, sizeof(double) * S->Np, S->s));
      }
      NC_CHECK_HIP(hipStreamSynchronize(S->s));
    } catch (...) {
      delete S;
      throw;
    }
    return S;
  }
}

// Re-initialise integrator state from the device copy u0 (time t0, norm beta0).
static void solver_reinit(nc_solver_s* S) {
  {
    const unsigned gc = std::min<unsigned>(4096, grid_for(S->Np, kBlock));
    k_copy_d<<<gc, kBlock, 0, S->s>>>(S->u0.p, S->u.p, S->Np);
    S->t = S->t0;
    S->have_state = true;
    S->fsal_valid = false;
    S->started = false;
    S->h = 0.0;
    S->st = nc_solve_stats{};
    if (S->method == NC_METHOD_KRYLOV) {
      S->beta = S->beta0;
      S->beta_pending = false;
      const KLayout L{S->m};
      const double sc0 = S->beta > 0 ? 1.0 / S->beta : 0.0;
      NC_CHECK_HIP(hipMemcpyAsync(S->kbuf.p + L.sc(0), &sc0, sizeof(double), hipMemcpyHostToDevice, S->s));
      const unsigned g = grid_for(S->Np, kBlock);
      if (S->A->precision == NC_FP32)
        k_to_vt<float><<<std::min(g, 4096u), kBlock, 0, S->s>>>(S->u.p, reinterpret_cast<float*>(S->V.p), S->Np);
      else
        k_to_vt<double><<<std::min(g, 4096u), kBlock, 0, S->s>>>(S->u.p, reinterpret_cast<double*>(S->V.p), S->Np);
    }
    NC_CHECK_HIP(hipGetLastError());
  }
}

static void solver_set_state_impl(nc_solver_s* S, double t0, const double* u) {
  NC_REQUIRE(S && u, NC_ERR_ARG, "null argument");
  if (!S->u0.p) {
    S->u0.alloc(S->Np);
    NC_CHECK_HIP(hipMemsetAsync(S->u0.p, 0, sizeof(double) * S->Np, S->s));
  }
  copy_from_host(S, S->u0.p, u);
  double b2 = 0.0;
  for (int64_t i = 0; i < S->N; ++i) b2 += u[i] * u[i];
  S->beta0 = std::sqrt(b2);
  S->t0 = t0;
  solver_reinit(S);
  NC_CHECK_HIP(hipStreamSynchronize(S->s));
}

static void solver_run_impl(nc_solver_s* S, int32_t n_out, const double* t_out) {
  {
    NC_REQUIRE(S && S->have_state, NC_ERR_ARG, "solver has no state; call nc_solver_set_state first");
    NC_REQUIRE(n_out >= 0 && (n_out == 0 || t_out), NC_ERR_ARG, "bad output times");
    for (int k = 0; k < n_out; ++k) {
      NC_REQUIRE(std::isfinite(t_out[k]) && t_out[k] >= S->t - 1e-12 * std::max(1.0, std::fabs(S->t)), NC_ERR_ARG,
                 "output times must be finite and not before the current time");
      NC_REQUIRE(k == 0 || t_out[k] >= t_out[k - 1], NC_ERR_ARG, "output times must be ascending");
    }
    if ((size_t)n_out * S->Np > S->snaps.n) S->snaps.alloc((size_t)std::max(n_out, 1) * S->Np);
    S->n_snaps = n_out;
    const unsigned g = std::min<unsigned>(4096, grid_for(S->Np, kBlock));
    for (int k = 0; k < n_out; ++k) {
      if (S->method == NC_METHOD_KRYLOV)
        krylov_advance(S, t_out[k]);
      else
        rk_advance(S, t_out[k]);
      k_copy_d<<<g, kBlock, 0, S->s>>>(S->u.p, S->snaps.p + (size_t)k * S->Np, S->Np);
      NC_CHECK_HIP(hipGetLastError());
    }
    NC_CHECK_HIP(hipStreamSynchronize(S->s));
    matvec_harvest_timing(S->A);
  }
}

}  // namespace

extern "C" {

int nc_solve_opts_default(nc_solve_opts* o) {
  if (!o) return NC_ERR_ARG;
  std::memset(o, 0, sizeof(*o));
  o->method = NC_METHOD_AUTO;
  o->krylov_m = 30;
  o->krylov_tol = 1e-12;
  o->rtol = 1e-8;
  o->atol = 1e-14;
  return NC_OK;
}

int nc_solver_create(nc_matrix A, const nc_solve_opts* opts, nc_solver* out) {
  return guard([&] {
    NC_REQUIRE(out, NC_ERR_ARG, "null argument");
    *out = solver_create_impl(A, opts);
  });
}

int nc_solver_destroy(nc_solver S) {
  return guard([&] { delete S; });
}

int nc_solver_set_state(nc_solver S, double t0, const double* u) {
  return guard([&] { solver_set_state_impl(S, t0, u); });
}

int nc_solver_restart(nc_solver S) {
  return guard([&] {
    NC_REQUIRE(S && S->u0.p, NC_ERR_ARG, "solver has no state; call nc_solver_set_state first");
    solver_reinit(S);
  });
}

int nc_solver_run(nc_solver S, int32_t n_out, const double* t_out) {
  return guard([&] { solver_run_impl(S, n_out, t_out); });
}

int nc_solver_get_snapshot(nc_solver S, int32_t k, double* u) {
  return guard([&] {
    NC_REQUIRE(S && u && k >= 0 && k < S->n_snaps, NC_ERR_ARG, "snapshot index out of range");
    copy_to_host(S, u, S->snaps.p + (size_t)k * S->Np);
  });
}

int nc_solver_get_fspvec(nc_solver S, int32_t k, int32_t block, nc_space space, nc_fspvec* p, double* sinks) {
  return guard([&] {
    NC_REQUIRE(S && space && p && k >= 0 && k < S->n_snaps, NC_ERR_ARG, "snapshot index out of range");
    NC_REQUIRE(block >= 0 && block < S->blocks, NC_ERR_ARG, "block index out of range");
    NC_REQUIRE(space->n == S->A->n && space->nr == S->A->nr, NC_ERR_ARG,
               "the state space does not match the solver's generator");
    if (!S->snap_states || S->snap_space_uid != space->uid || S->snap_version != space->version) {
      S->snap_states = share_states(space->states.p, space->n, space->ns, space->stream.s);
      NC_CHECK_HIP(hipStreamSynchronize(space->stream.s));
      S->snap_space_uid = space->uid;
      S->snap_version = space->version;
    }
    const double* base = S->snaps.p + (size_t)k * S->Np + (size_t)block * S->Npb;
    *p = fspvec_from_device(space->ns, space->n, S->snap_states, base, S->s);
    if (sinks)
      NC_CHECK_HIP(hipMemcpy(sinks, base + S->A->n, sizeof(double) * S->A->nr, hipMemcpyDeviceToHost));
  });
}

int nc_solver_get_state(nc_solver S, double* t, double* u) {
  return guard([&] {
    NC_REQUIRE(S, NC_ERR_ARG, "null solver");
    if (t) *t = S->t;
    if (u) copy_to_host(S, u, S->u.p);
  });
}

int nc_solver_stats(nc_solver S, nc_solve_stats* st) {
  return guard([&] {
    NC_REQUIRE(S && st, NC_ERR_ARG, "null argument");
    *st = S->st;
  });
}

static void solve_fixed_impl(nc_matrix_s* A, nc_sensmatrix_s* SA, const nc_solve_opts* opts, const double* u0,
                             int32_t n_out, const double* t_out, double* u_out, nc_solve_stats* st) {
  NC_REQUIRE(A && u0 && (n_out == 0 || (t_out && u_out)), NC_ERR_ARG, "null argument");
  nc_solver_s* S = solver_create_impl(A, opts, SA);
  struct Guard {
    nc_solver_s* S;
    ~Guard() { delete S; }
  } gd{S};
  const double t0 = n_out ? t_out[0] : 0.0;
  solver_set_state_impl(S, t0, u0);
  solver_run_impl(S, n_out, t_out);
  for (int k = 0; k < n_out; ++k) copy_to_host(S, u_out + (size_t)k * S->N, S->snaps.p + (size_t)k * S->Np);
  if (st) nc_solver_stats(S, st);
}

int nc_solve_fixed(nc_matrix A, const nc_solve_opts* opts, const double* u0, int32_t n_out, const double* t_out,
                   double* u_out, nc_solve_stats* st) {
  return guard([&] { solve_fixed_impl(A, nullptr, opts, u0, n_out, t_out, u_out, st); });
}

int nc_solver_create_sens(nc_sensmatrix SA, const nc_solve_opts* opts, nc_solver* out) {
  return guard([&] {
    NC_REQUIRE(SA && out, NC_ERR_ARG, "null argument");
    *out = solver_create_impl(SA->A, opts, SA);
  });
}

int nc_solve_sens_fixed(nc_sensmatrix SA, const nc_solve_opts* opts, const double* u0, int32_t n_out,
                        const double* t_out, double* u_out, nc_solve_stats* st) {
  return guard([&] {
    NC_REQUIRE(SA, NC_ERR_ARG, "null argument");
    solve_fixed_impl(SA->A, SA, opts, u0, n_out, t_out, u_out, st);
  });
}

}  // extern "C"

namespace nc {

nc_solver_s* solver_create(nc_matrix_s* A, const nc_solve_opts* opts, nc_sensmatrix_s* SA) {
  return solver_create_impl(A, opts, SA);
}

// Set the state from a device vector of the true length N, blocks contiguous (t0 is its
// time).
double solver_step_hint(const nc_solver_s* S) {
  if (S->method == NC_METHOD_KRYLOV) return S->started ? S->t_new : 0.0;
  return S->h;
}

void solver_seed_step(nc_solver_s* S, double h) {
  if (!(h > 0.0) || !std::isfinite(h)) return;
  if (S->method == NC_METHOD_KRYLOV) {
    S->anorm = std::max(S->SA ? S->SA->norm1 : S->A->norm1, 1e-300);
    S->t_new = h;
    S->started = true;
  } else {
    S->h = h;
  }
}

void solver_set_state_device(nc_solver_s* S, double t0, const double* u_dev) {
  if (!S->u0.p) {
    S->u0.alloc(S->Np);
    NC_CHECK_HIP(hipMemsetAsync(S->u0.p, 0, sizeof(double) * S->Np, S->s));
  }
  // u_dev holds the blocks contiguously (stride Nb); the solver pads each to Npb
  NC_CHECK_HIP(hipMemcpy2DAsync(S->u0.p, sizeof(double) * S->Npb, u_dev, sizeof(double) * S->Nb,
                                sizeof(double) * S->Nb, S->blocks, hipMemcpyDeviceToDevice, S->s));
  const KLayout L{S->m};
  VecArgs<double> a{};
  a.N = S->Np;
  a.nv = 0;
  a.xin = S->u0.p;
  a.partials = S->partials.p;
  a.kb = S->kbuf.p + L.misc(5);
  a.L = L;
  a.finish = FIN_SUMS;
  launch_vec(a, S, S->s);
  double b2 = 0.0;
  NC_CHECK_HIP(hipMemcpyAsync(&b2, S->kbuf.p + L.misc(5), sizeof(double), hipMemcpyDeviceToHost, S->s));
  NC_CHECK_HIP(hipStreamSynchronize(S->s));
  S->beta0 = std::sqrt(b2);
  S->t0 = t0;
  solver_reinit(S);
  NC_CHECK_HIP(hipStreamSynchronize(S->s));
}

}  // namespace nc
