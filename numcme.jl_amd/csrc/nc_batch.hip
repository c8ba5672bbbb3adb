// Batched FSP solves: many independent generators of uneven size integrated together.
//
// Layout: every instance's slot-major generator is copied into one concatenated ELL
// (ld = total rows). Instance i owns rows [row_off[i], row_off[i] + rows_pad[i]), padded
// to whole 256-row granules; gran_inst[granule] names the owner. A thread works on R
// consecutive rows (R = 1, 2 or 4, chosen by nc_batch_autotune), which never straddle a
// granule, so every thread has one owner instance. Rows n_i .. n_i + nr_i - 1 of an
// instance are its sinks (written by one workgroup per (instance, reaction) segment);
// padding rows stay zero. Vectors use the same row space.
//
// Integrator: Dormand-Prince 5(4) with per-instance t, h and error control, all of it on
// the device. One iteration is
//   accept/snapshot of the previous step fused with stage 1, five more stages with their
//   batched mat-vecs, the error partials, their per-instance fold, and the controller
// The controller (one workgroup) decides every instance's accept, snapshot and next step
// size, and lists the granules of the instances that have work in the next iteration;
// every row kernel walks that list with a fixed grid, so an iteration costs what its
// active instances cost (the large instances run longest: late iterations touch only
// them). Iterations are captured in chunks as one HIP graph and replayed with no host
// decision in between; the host reads one counter per chunk (the number of instances
// still running), with the next chunk already queued.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <memory>
#include <mutex>
#include <string>
#include <thread>

#include "nc_matrix.h"

namespace nc {
namespace {

constexpr int kBlock = 256;
constexpr int kGran = 256;       // rows of one ownership granule (instances are padded to it)
constexpr int kCtlBlock = 1024;  // controller workgroup
constexpr unsigned kMaxRowBlocks = 2048;  // fixed grid of the list-driven row kernels
constexpr int kChunk = 8;        // iterations per captured graph
constexpr int kBuildWorkers = 4;  // host threads of nc_batch_create_models

template <typename T, int R>
struct alignas(sizeof(T) * R) Pack {
  T v[R];
};

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__constant__ double c_a[7][6] = {
    {0, 0, 0, 0, 0, 0},
    {1.0 / 5, 0, 0, 0, 0, 0},
    {3.0 / 40, 9.0 / 40, 0, 0, 0, 0},
    {44.0 / 45, -56.0 / 15, 32.0 / 9, 0, 0, 0},
    {19372.0 / 6561, -25360.0 / 2187, 64448.0 / 6561, -212.0 / 729, 0, 0},
    {9017.0 / 3168, -355.0 / 33, 46732.0 / 5247, 49.0 / 176, -5103.0 / 18656, 0},
    {35.0 / 384, 0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784, 11.0 / 84}};
__constant__ double c_e[7] = {71.0 / 57600, 0, -71.0 / 16695, 71.0 / 1920, -17253.0 / 339200, 22.0 / 525, -1.0 / 40};

// Solve settings read by the captured kernels from device memory, so that one captured
// chunk serves every solve.
struct Params {
  double atol, rtol, h0;
  int64_t max_steps;  // steps + rejections per instance, 0 = unlimited
  int32_t n_out;
};

struct Dev {
  const int32_t* gran_inst;  // owner instance of each granule
  const int64_t* row_off;    // first row of each instance
  const int64_t* nstates;    // states of each instance
  const double* h;           // step size of the current iteration (0: inactive)
  const int32_t* accept;     // 1: accept the last step
  const int32_t* snap;       // snapshot slot to fill after accepting, or -1
  const int32_t* agran;      // granules of the instances with work in this iteration
  const uint32_t* n_agran;   // their count
};

// Work unit u of a list-driven kernel covers list entries u R .. u R + R - 1 (R granules,
// kBlock R rows); thread t takes R consecutive rows of entry u R + t R / kGran. Returns
// false for a unit entry past the end of the list.
template <int R>
__device__ inline bool unit_rows(const Dev& d, uint32_t n_list, int64_t u, int64_t& n0) {
  const int64_t j = u * R + ((int64_t)threadIdx.x * R) / kGran;
  if (j >= (int64_t)n_list) return false;
  n0 = (int64_t)d.agran[j] * kGran + ((int64_t)threadIdx.x * R) % kGran;
  return true;
}

// y_s = u + h_i sum_{j<s} a_sj k_j over the instance's rows: its VT copy, and at s = 6
// (y_{n+1}, read by the error partials and the accept) its FP64 value
// (the stage is a template argument so that its STAGE loads of k go out together)
template <typename VT, int R, int STAGE>
__global__ __launch_bounds__(kBlock) void k_bstage(Dev d, int64_t ld, const double* u, const double* k,
                                                   double* ytmp, VT* vin) {
  const uint32_t nl = *d.n_agran;
  const int64_t units = ((int64_t)nl + R - 1) / R;
  for (int64_t un = blockIdx.x; un < units; un += gridDim.x) {
    int64_t n0;
    if (!unit_rows<R>(d, nl, un, n0)) continue;
    const double h = d.h[d.gran_inst[n0 / kGran]];
    if (h == 0.0) continue;
    Pack<double, R> y = *reinterpret_cast<const Pack<double, R>*>(u + n0);
    // a_61 = 0: the last stage does not read k1
    Pack<double, R> kj[STAGE > 0 ? STAGE : 1];
#pragma unroll
    for (int j = 0; j < STAGE; ++j)
      if (!(STAGE == 6 && j == 1)) kj[j] = *reinterpret_cast<const Pack<double, R>*>(k + (int64_t)j * ld + n0);
    double acc[R] = {};
#pragma unroll
    for (int j = 0; j < STAGE; ++j) {
      if (STAGE == 6 && j == 1) continue;
#pragma unroll
      for (int r = 0; r < R; ++r) acc[r] += c_a[STAGE][j] * kj[j].v[r];
    }
    Pack<VT, R> o;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      y.v[r] += h * acc[r];
      o.v[r] = (VT)y.v[r];
    }
    if constexpr (STAGE == 6) *reinterpret_cast<Pack<double, R>*>(ytmp + n0) = y;
    *reinterpret_cast<Pack<VT, R>*>(vin + n0) = o;
  }
}

// out = A v over the state rows of active instances (R rows per thread). With the slot
// count KT as a template argument every slot's column and value load goes out in one
// batch and every gather in a second (KT = 0: runtime count K, one slot at a time).
// The nseg sink segments (one per instance and reaction with sink entries) come first in
// the grid, one workgroup each, so that their reductions overlap the row blocks.
struct SinkSegs {
  int32_t nseg;
  const int32_t* inst;  // owner instance
  const int64_t* row;   // output row of the sink
  const int64_t* off;   // nseg + 1: entry ranges
  const int32_t* src;   // source state rows of the entries
};

// DM 1: instances whose every slot column is the row plus a constant (dia_inst[inst] = 1,
// a space in lattice order; the offsets are dia_off[inst * Kb + s]) compute those columns,
// clamped to the instance's states (rows whose value is 0 may point anywhere), instead of
// loading them; the other instances load theirs. A wave's rows lie in one granule, so the
// instance, its kind and its offsets are uniform and the choice is a scalar branch.
template <typename VT, int R, int KT, int DM>
__global__ __launch_bounds__(kBlock) void k_bmatvec(Dev d, int64_t ld, int K, const double* diag, const int32_t* col,
                                                    const VT* val, const VT* v, double* out, SinkSegs sg,
                                                    const VT* sval, const int32_t* dia_off, const int32_t* dia_inst,
                                                    int Kb) {
  if ((int)blockIdx.x < sg.nseg) {
    // sink row of one (instance, reaction) segment: one workgroup, deterministic tree
    const int seg = blockIdx.x;
    if (d.h[sg.inst[seg]] == 0.0) return;
    double acc = 0.0;
    for (int64_t e = sg.off[seg] + threadIdx.x; e < sg.off[seg + 1]; e += kBlock)
      acc += (double)sval[e] * (double)v[sg.src[e]];
    __shared__ double part[kBlock / 64];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[sg.row[seg]] = part[0] + part[1] + part[2] + part[3];
    return;
  }
  const uint32_t nl = *d.n_agran;
  const int64_t units = ((int64_t)nl + R - 1) / R;
  for (int64_t un = (int64_t)blockIdx.x - sg.nseg; un < units; un += (int64_t)gridDim.x - sg.nseg) {
    int64_t n0;
    if (!unit_rows<R>(d, nl, un, n0)) continue;
    const int inst = __builtin_amdgcn_readfirstlane(d.gran_inst[n0 / kGran]);
    if (d.h[inst] == 0.0) continue;
    const int64_t nbeg = d.row_off[inst], nend = nbeg + d.nstates[inst];
    if (n0 >= nend) continue;
    const Pack<double, R> dg = *reinterpret_cast<const Pack<double, R>*>(diag + n0);
    const Pack<VT, R> vi = *reinterpret_cast<const Pack<VT, R>*>(v + n0);
    double acc[R];
    if constexpr (KT > 0) {
      Pack<int32_t, R> c[KT];
      Pack<VT, R> w[KT];
      if (DM == 1 && dia_inst[inst]) {
#pragma unroll
        for (int s = 0; s < KT; ++s) {
          const int64_t off = dia_off[(int64_t)inst * Kb + s];
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const int64_t j = n0 + r + off;
            c[s].v[r] = (int32_t)(j < nbeg ? nbeg : (j >= nend ? nend - 1 : j));
          }
          w[s] = *reinterpret_cast<const Pack<VT, R>*>(val + (int64_t)s * ld + n0);
        }
      } else {
#pragma unroll
        for (int s = 0; s < KT; ++s) {
          c[s] = *reinterpret_cast<const Pack<int32_t, R>*>(col + (int64_t)s * ld + n0);
          w[s] = *reinterpret_cast<const Pack<VT, R>*>(val + (int64_t)s * ld + n0);
        }
      }
      double gv[KT][R];
#pragma unroll
      for (int s = 0; s < KT; ++s) {
#pragma unroll
        for (int r = 0; r < R; ++r) gv[s][r] = (double)v[c[s].v[r]];
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int r = 0; r < R; ++r) acc[r] = dg.v[r] * (double)vi.v[r];
#pragma unroll
      for (int s = 0; s < KT; ++s) {
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] += (double)w[s].v[r] * gv[s][r];
      }
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) acc[r] = dg.v[r] * (double)vi.v[r];
      for (int s = 0; s < K; ++s) {
        const Pack<int32_t, R> c = *reinterpret_cast<const Pack<int32_t, R>*>(col + (int64_t)s * ld + n0);
        const Pack<VT, R> w = *reinterpret_cast<const Pack<VT, R>*>(val + (int64_t)s * ld + n0);
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] += (double)w.v[r] * (double)v[c.v[r]];
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (n0 + r < nend) out[n0 + r] = acc[r];
  }
}

// Error partials per granule: sum ((h sum_j e_j k_j) / (atol + rtol max(|u|, |y|)))^2.
// A unit covers R list entries; entry j of the unit is waves [j W, (j + 1) W), W = 4 / R.
template <int R>
__global__ __launch_bounds__(kBlock) void k_berr(Dev d, int64_t ld, const Params* par, const double* u,
                                                 const double* y, const double* k, double* partials) {
  constexpr int W = (kBlock / 64) / R;
  __shared__ double part[kBlock / 64];
  const double atol = par->atol, rtol = par->rtol;
  const uint32_t nl = *d.n_agran;
  const int64_t units = ((int64_t)nl + R - 1) / R;
  for (int64_t un = blockIdx.x; un < units; un += gridDim.x) {
    double acc = 0.0;
    int64_t n0 = 0;
    const bool ok = unit_rows<R>(d, nl, un, n0);
    const double h = ok ? d.h[d.gran_inst[n0 / kGran]] : 0.0;
    if (h != 0.0) {
      const Pack<double, R> uu = *reinterpret_cast<const Pack<double, R>*>(u + n0);
      const Pack<double, R> yy = *reinterpret_cast<const Pack<double, R>*>(y + n0);
      double e[R] = {};
      // e_1 = 0 in Dormand-Prince 5(4): k1 is not read (a compile-time skip, so that the
      // six loads go out together)
      Pack<double, R> kj[7];
#pragma unroll
      for (int j = 0; j < 7; ++j)
        if (j != 1) kj[j] = *reinterpret_cast<const Pack<double, R>*>(k + (int64_t)j * ld + n0);
#pragma unroll
      for (int j = 0; j < 7; ++j) {
        if (j == 1) continue;
#pragma unroll
        for (int r = 0; r < R; ++r) e[r] += c_e[j] * kj[j].v[r];
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double sc = atol + rtol * fmax(fabs(uu.v[r]), fabs(yy.v[r]));
        const double q = h * e[r] / sc;
        acc += q * q;
      }
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if ((int)threadIdx.x < R) {
      const int64_t j = un * R + threadIdx.x;
      if (j < (int64_t)nl) {
        double sum = 0.0;
        for (int w = 0; w < W; ++w) sum += part[threadIdx.x * W + w];
        partials[d.agran[j]] = sum;
      }
    }
    __syncthreads();
  }
}

// Per-instance fold of the granule partials (granules of an instance are contiguous): one
// wavefront per instance, lane l adding granules l, l + 64, ... in order, then a fixed
// shuffle tree (deterministic)
__global__ __launch_bounds__(kBlock) void k_bfold(int n_inst, const int64_t* gran_start, const double* h,
                                                  const double* partials, double* err) {
  const int i = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n_inst) return;  // whole wavefronts leave together
  if (h[i] == 0.0) return;  // not stepped in this iteration: its partials are stale
  double sum = 0.0;
  for (int64_t g = gran_start[i] + lane; g < gran_start[i + 1]; g += 64) sum += partials[g];
  sum = wave_sum(sum);
  if (lane == 0) err[i] = sum;
}

// Accept the last step where asked (u <- y, k0 <- k6) and fill requested snapshots.
template <int R>
__global__ __launch_bounds__(kBlock) void k_baccept(Dev d, int64_t ld, double* u, const double* y, double* k,
                                                    double* snaps) {
  const uint32_t nl = *d.n_agran;
  const int64_t units = ((int64_t)nl + R - 1) / R;
  for (int64_t un = blockIdx.x; un < units; un += gridDim.x) {
    int64_t n0;
    if (!unit_rows<R>(d, nl, un, n0)) continue;
    const int inst = d.gran_inst[n0 / kGran];
    if (d.accept[inst]) {
      *reinterpret_cast<Pack<double, R>*>(u + n0) = *reinterpret_cast<const Pack<double, R>*>(y + n0);
      *reinterpret_cast<Pack<double, R>*>(k + n0) = *reinterpret_cast<const Pack<double, R>*>(k + 6 * ld + n0);
    }
    const int sl = d.snap[inst];
    if (sl >= 0)
      *reinterpret_cast<Pack<double, R>*>(snaps + (int64_t)sl * ld + n0) =
          *reinterpret_cast<const Pack<double, R>*>(u + n0);
  }
}

// k_baccept followed by stage 1 in one pass: the accepted (or kept) u and k0 stay in
// registers for the stage, instead of being written by one kernel and read by the next.
template <typename VT, int R>
__global__ __launch_bounds__(kBlock) void k_baccept_stage1(Dev d, int64_t ld, double* u, const double* y, double* k,
                                                           double* snaps, VT* vin) {
  const uint32_t nl = *d.n_agran;
  const int64_t units = ((int64_t)nl + R - 1) / R;
  for (int64_t un = blockIdx.x; un < units; un += gridDim.x) {
    int64_t n0;
    if (!unit_rows<R>(d, nl, un, n0)) continue;
    const int inst = d.gran_inst[n0 / kGran];
    Pack<double, R> uu, k0;
    if (d.accept[inst]) {
      uu = *reinterpret_cast<const Pack<double, R>*>(y + n0);
      k0 = *reinterpret_cast<const Pack<double, R>*>(k + 6 * ld + n0);
      *reinterpret_cast<Pack<double, R>*>(u + n0) = uu;
      *reinterpret_cast<Pack<double, R>*>(k + n0) = k0;
    } else {
      uu = *reinterpret_cast<const Pack<double, R>*>(u + n0);
      k0 = *reinterpret_cast<const Pack<double, R>*>(k + n0);
    }
    const int sl = d.snap[inst];
    if (sl >= 0) *reinterpret_cast<Pack<double, R>*>(snaps + (int64_t)sl * ld + n0) = uu;
    const double h = d.h[inst];
    if (h == 0.0) continue;
    Pack<VT, R> o;
#pragma unroll
    for (int r = 0; r < R; ++r) o.v[r] = (VT)(uu.v[r] + h * (c_a[1][0] * k0.v[r]));
    *reinterpret_cast<Pack<VT, R>*>(vin + n0) = o;
  }
}

// ---- the step-size controller ----------------------------------------------------------
// Per-instance state of the Dormand-Prince controller, in device memory.
struct Ctl {
  int n;
  const Params* par;
  const double* t_out;
  const double* err;        // folded error sums of the iteration just run
  const int64_t* nrows;     // error norm length per instance
  const int64_t* gran_start;
  double* t;                // current time
  double* hc;               // controller step size
  double* htry;             // step of the next iteration (0: inactive); Dev::h
  int32_t* next;            // next output index
  int32_t* done;
  int32_t* clamped;         // htry was cut to hit an output time
  int32_t* accept;          // Dev::accept
  int32_t* snap;            // Dev::snap
  int64_t* steps;
  int64_t* rejected;
  int32_t* status;          // 0 ok, 1 non-finite error, 2 step underflow, 3 max_steps
  uint32_t* off;            // n + 1 list offsets (scratch)
  int32_t* agran;           // Dev::agran of the next iteration
  uint32_t* n_agran;        // Dev::n_agran
  uint32_t* counters;       // [0] instances still running, [1] iterations with work
};

__device__ void control_one(const Ctl& c, int i, bool init, int n_out, double h0, int64_t max_steps) {
  if (init) {
    c.t[i] = n_out ? c.t_out[0] : 0.0;
    c.next[i] = 1;
    c.done[i] = n_out <= 1 ? 1 : 0;
    c.steps[i] = 0;
    c.rejected[i] = 0;
    c.status[i] = 0;
    c.accept[i] = 0;
    c.snap[i] = n_out > 0 ? 0 : -1;  // the start is the first snapshot
    const double tend = n_out ? c.t_out[n_out - 1] : 0.0;
    // first step: a fraction of the span (the controller adapts it within a few steps)
    c.hc[i] = h0 > 0 ? h0 : 1e-3 * fmax(1e-12, tend - c.t[i]);
  } else if (c.htry[i] != 0.0) {
    const double hs = c.htry[i];
    const double err = sqrt(c.err[i] / (double)c.nrows[i]);
    c.accept[i] = 0;
    c.snap[i] = -1;
    if (!isfinite(err)) {
      c.status[i] = 1;
      c.done[i] = 1;
    } else {
      double fac = err > 0 ? 0.9 * pow(err, -0.2) : 5.0;
      double h;
      if (err <= 1.0) {
        c.accept[i] = 1;
        c.steps[i]++;
        double t = c.t[i] + hs;
        const double target = c.t_out[c.next[i]];
        if (t >= target - 1e-13 * fmax(1.0, fabs(target))) {
          t = target;
          c.snap[i] = c.next[i];
          if (++c.next[i] >= n_out) c.done[i] = 1;
        }
        c.t[i] = t;
        fac = fmin(5.0, fmax(0.2, fac));
        h = (!c.clamped[i] || fac < 1.0) ? hs * fac : c.hc[i];
      } else {
        c.rejected[i]++;
        h = hs * fmax(0.2, fac);
      }
      c.hc[i] = h;
      if (!c.done[i] && !(h > 1e-14 * fmax(1.0, fabs(c.t[i])))) c.status[i] = 2, c.done[i] = 1;
      if (!c.done[i] && max_steps && c.steps[i] + c.rejected[i] >= max_steps) c.status[i] = 3, c.done[i] = 1;
    }
  } else {
    c.accept[i] = 0;
    c.snap[i] = -1;
  }
  // the step of the next iteration, cut to the next output time
  if (c.done[i]) {
    c.htry[i] = 0.0;
  } else {
    const double target = c.t_out[c.next[i]];
    const double hh = fmin(c.hc[i], target - c.t[i]);
    c.clamped[i] = hh < c.hc[i];
    c.htry[i] = hh;
  }
}

// Block-wide exclusive scan of one value per thread (kCtlBlock threads); returns the
// thread's offset and the block total.
__device__ uint32_t block_scan(uint32_t v, uint32_t* sh, uint32_t& total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) sh[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int k = 0; k < kCtlBlock / 64; ++k) {
      const uint32_t s = sh[k];
      sh[k] = run;
      run += s;
    }
    sh[kCtlBlock / 64] = run;
  }
  __syncthreads();
  const uint32_t res = sh[w] + x - v;
  total = sh[kCtlBlock / 64];
  __syncthreads();
  return res;
}

// One workgroup: every instance's controller step, then the granule list of the instances
// with work in the next iteration (a step to take, or an accept / snapshot to apply).
__global__ __launch_bounds__(kCtlBlock) void k_bcontrol(Ctl c, int init) {
  __shared__ uint32_t sh[kCtlBlock / 64 + 1];
  __shared__ int any;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  const int n_out = c.par->n_out;
  const double h0 = c.par->h0;
  const int64_t max_steps = c.par->max_steps;
  uint32_t running = 0;
  for (int i = threadIdx.x; i < c.n; i += kCtlBlock) {
    if (!init && c.htry[i] != 0.0) atomicOr(&any, 1);
    control_one(c, i, init != 0, n_out, h0, max_steps);
    running += c.htry[i] != 0.0;
  }
  // instance offsets into the list, in instance order
  uint32_t carry = 0;
  for (int base = 0; base < c.n; base += kCtlBlock) {
    const int i = base + threadIdx.x;
    uint32_t cnt = 0;
    if (i < c.n && (c.htry[i] != 0.0 || c.accept[i] || c.snap[i] >= 0))
      cnt = (uint32_t)(c.gran_start[i + 1] - c.gran_start[i]);
    uint32_t tot;
    const uint32_t o = block_scan(cnt, sh, tot);
    if (i < c.n) c.off[i] = carry + o;
    carry += tot;
  }
  if (threadIdx.x == 0) c.off[c.n] = carry;
  // running count: one wave sum per wave, then atomics (ordering-free: integers)
  __shared__ uint32_t nrun;
  if (threadIdx.x == 0) nrun = 0;
  __syncthreads();
  atomicAdd(&nrun, running);
  __syncthreads();
  // the list: instance i's granules at entries [off[i], off[i + 1]), written by one wave
  // per instance (instances dealt to the waves in turn), 64 consecutive entries at a time
  const uint32_t total = c.off[c.n];
  {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int i = wave; i < c.n; i += kCtlBlock / 64) {
      const uint32_t o0 = c.off[i], o1 = c.off[i + 1];
      const int64_t g0 = c.gran_start[i];
      for (uint32_t w = o0 + lane; w < o1; w += 64) c.agran[w] = (int32_t)(g0 + (w - o0));
    }
  }
  if (threadIdx.x == 0) {
    *c.n_agran = total;
    c.counters[0] = nrun;
    if (any) c.counters[1]++;
  }
}

// Copy one instance's slot-major ELL into the concatenated arrays.
template <typename VT>
__global__ void k_bcopy(int64_t n, int64_t ld_src, int K_src, int K, int64_t off, int64_t ld, const double* diag_s,
                        const int32_t* col_s, const VT* val_s, double* diag, int32_t* col, VT* val) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  diag[off + i] = diag_s[i];
  for (int s = 0; s < K; ++s) {
    if (s < K_src) {
      col[s * ld + off + i] = col_s[s * ld_src + i] + (int32_t)off;
      val[s * ld + off + i] = val_s[s * ld_src + i];
    } else {
      col[s * ld + off + i] = (int32_t)(off + i);
      val[s * ld + off + i] = (VT)0;
    }
  }
}

__global__ void k_badd_off(int64_t n, const int32_t* src_s, int64_t off, int32_t* src) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) src[i] = src_s[i] + (int32_t)off;
}

// Every instance listed, with step h (the autotuner's frozen iterations).
__global__ void k_bfreeze(int n, double h, const int64_t* gran_start, double* htry, int32_t* accept, int32_t* snap,
                          int32_t* agran, uint32_t* n_agran) {
  const int64_t total = gran_start[n];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) agran[g] = (int32_t)g;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) htry[i] = h, accept[i] = 0, snap[i] = -1;
  if (i == 0) *n_agran = (uint32_t)total;
}

// Snapshot rows of every instance packed back to back (instance i at pack_off[i]), one
// workgroup per instance: the layout of the solve's initial vectors.
__global__ __launch_bounds__(kBlock) void k_bpack(const double* __restrict__ snap, const int64_t* __restrict__ row_off,
                                                  const int64_t* __restrict__ nrows,
                                                  const int64_t* __restrict__ pack_off, double* __restrict__ out) {
  const int i = blockIdx.x;
  const double* src = snap + row_off[i];
  double* dst = out + pack_off[i];
  const int64_t n = nrows[i];
  for (int64_t r = threadIdx.x; r < n; r += kBlock) dst[r] = src[r];
}

}  // namespace
}  // namespace nc

struct nc_batch_s {
  int n = 0, precision = NC_FP64;
  size_t vs = 8;
  int K = 0;
  int64_t ld = 0, grans = 0;
  int rows_per_thread = 4;  // R of the kernels (1, 2 or 4)
  std::vector<int64_t> row_off, nstates, nrows, gran_start;
  nc::Stream stream;
  nc::DevBuf<double> diag;
  nc::DevBuf<int32_t> col;
  nc::DevBuf<unsigned char> val;
  // instances whose slot columns are the row plus a constant (dia_inst[i] = 1) and their
  // offsets (n x K): the mat-vec computes those columns instead of loading them
  bool any_dia = false;
  nc::DevBuf<int32_t> dia_off, dia_inst;
  int nseg = 0;
  nc::DevBuf<int32_t> seg_inst, sink_src;
  nc::DevBuf<int64_t> seg_row, seg_off;
  nc::DevBuf<unsigned char> sink_val;
  nc::DevBuf<int32_t> gran_inst;
  nc::DevBuf<int64_t> d_row_off, d_nstates, d_gran_start, d_nrows;
  // integrator state
  nc::DevBuf<double> u, ytmp, k, partials, err, snaps;
  nc::DevBuf<unsigned char> vin;
  // controller state (per instance) and the active granule list
  nc::DevBuf<double> c_t, c_hc, d_h, t_out;
  nc::DevBuf<int32_t> c_next, c_done, c_clamped, d_accept, d_snap, c_status;
  nc::DevBuf<int64_t> c_steps, c_rej;
  nc::DevBuf<uint32_t> c_off, n_agran, counters;
  nc::DevBuf<int32_t> agran;
  nc::DevBuf<nc::Params> par;
  uint32_t* cnt_pin = nullptr;  // pinned copies of the running counter, one per chunk in flight
  int n_out = 0;
  int n_out_cap = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_R = 0;
  ~nc_batch_s() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cnt_pin) nc::pinned_free(cnt_pin);
  }
};

using namespace nc;

namespace {

template <typename VT>
void batch_build(nc_batch_s* B, const nc_matrix_s* const* mats) {
  hipStream_t s = B->stream.s;
  for (int i = 0; i < B->n; ++i) {
    const nc_matrix_s* A = mats[i];
    NC_CHECK_HIP(hipStreamSynchronize(A->stream.s));
    if (A->n)
      k_bcopy<VT><<<grid_for(A->n, kBlock), kBlock, 0, s>>>(
          A->n, A->ld, A->K, B->K, B->row_off[i], B->ld, A->diag.p, A->col.p, reinterpret_cast<const VT*>(A->val.p),
          B->diag.p, B->col.p, reinterpret_cast<VT*>(B->val.p));
    NC_CHECK_HIP(hipGetLastError());
  }
  // computed columns: the instances in lattice order (each of their slots a constant offset)
  {
    std::vector<int32_t> offs((size_t)B->n * std::max(B->K, 1), 0), flag(B->n, 0);
    for (int i = 0; i < B->n; ++i) {
      const nc_matrix_s* A = mats[i];
      const uint32_t need = A->K >= 32 ? 0xffffffffu : ((1u << A->K) - 1u);
      if (A->K > (int)A->dia_off.size() || (A->dia_mask & need) != need) continue;
      flag[i] = 1;
      B->any_dia = true;
      for (int k = 0; k < A->K; ++k) offs[(size_t)i * B->K + k] = A->dia_off[k];
    }
    B->dia_off.alloc(offs.size());
    B->dia_inst.alloc(flag.size());
    NC_CHECK_HIP(hipMemcpyAsync(B->dia_off.p, offs.data(), sizeof(int32_t) * offs.size(), hipMemcpyHostToDevice, s));
    NC_CHECK_HIP(hipMemcpyAsync(B->dia_inst.p, flag.data(), sizeof(int32_t) * flag.size(), hipMemcpyHostToDevice, s));
    NC_CHECK_HIP(hipStreamSynchronize(s));
  }
  // sink segments: (instance, reaction), entries concatenated
  std::vector<int32_t> seg_inst;
  std::vector<int64_t> seg_row, seg_off{0};
  int64_t total = 0;
  for (int i = 0; i < B->n; ++i)
    for (int r = 0; r < mats[i]->nr; ++r) {
      seg_inst.push_back(i);
      seg_row.push_back(B->row_off[i] + mats[i]->n + r);
      total += mats[i]->sink_off_h[r + 1] - mats[i]->sink_off_h[r];
      seg_off.push_back(total);
    }
  B->nseg = (int)seg_inst.size();
  B->seg_inst.alloc(std::max<size_t>(seg_inst.size(), 1));
  B->seg_row.alloc(std::max<size_t>(seg_row.size(), 1));
  B->seg_off.alloc(seg_off.size());
  B->sink_src.alloc(std::max<int64_t>(total, 1));
  B->sink_val.alloc(std::max<int64_t>(total, 1) * B->vs);
  NC_CHECK_HIP(hipMemcpyAsync(B->seg_inst.p, seg_inst.data(), sizeof(int32_t) * seg_inst.size(), hipMemcpyHostToDevice, s));
  NC_CHECK_HIP(hipMemcpyAsync(B->seg_row.p, seg_row.data(), sizeof(int64_t) * seg_row.size(), hipMemcpyHostToDevice, s));
  NC_CHECK_HIP(hipMemcpyAsync(B->seg_off.p, seg_off.data(), sizeof(int64_t) * seg_off.size(), hipMemcpyHostToDevice, s));
  int64_t e = 0;
  for (int i = 0; i < B->n; ++i) {
    const nc_matrix_s* A = mats[i];
    const int64_t m = A->n_sink_entries;
    if (m) {
      k_badd_off<<<grid_for(m, kBlock), kBlock, 0, s>>>(m, A->sink_src.p, B->row_off[i], B->sink_src.p + e);
      NC_CHECK_HIP(hipMemcpyAsync(B->sink_val.p + e * B->vs, A->sink_val.p, m * B->vs, hipMemcpyDeviceToDevice, s));
    }
    e += m;
  }
  NC_CHECK_HIP(hipGetLastError());
  NC_CHECK_HIP(hipStreamSynchronize(s));
}

nc_batch_s* batch_create(int n, const nc_matrix* mats) {
  NC_REQUIRE(n >= 1 && mats, NC_ERR_ARG, "empty batch");
  auto B = std::make_unique<nc_batch_s>();
  B->n = n;
  B->precision = mats[0]->precision;
  B->vs = mats[0]->vsize;
  int64_t off = 0;
  for (int i = 0; i < n; ++i) {
    const nc_matrix_s* A = mats[i];
    NC_REQUIRE(A, NC_ERR_ARG, "null matrix in batch");
    NC_REQUIRE(A->precision == B->precision, NC_ERR_ARG, "all matrices of a batch must share one precision");
    NC_REQUIRE(A->tv_r.empty(), NC_ERR_UNSUPPORTED, "batched solves take time-invariant generators");
    B->K = std::max(B->K, A->K);
    const int64_t rows = (A->n_rows + kGran - 1) / kGran * kGran;
    B->row_off.push_back(off);
    B->nstates.push_back(A->n);
    B->nrows.push_back(A->n_rows);
    off += std::max<int64_t>(rows, kGran);
  }
  NC_REQUIRE(off < (int64_t)0x7FFFFFF0, NC_ERR_OVERFLOW, "batch exceeds 2^31 rows");
  B->ld = off;
  B->grans = off / kGran;
  std::vector<int32_t> gran_inst(B->grans);
  B->gran_start.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    const int64_t g0 = B->row_off[i] / kGran, g1 = (i + 1 < n ? B->row_off[i + 1] : off) / kGran;
    B->gran_start[i] = g0;
    for (int64_t g = g0; g < g1; ++g) gran_inst[g] = i;
  }
  B->gran_start[n] = B->grans;
  hipStream_t s = B->stream.s;
  const int64_t ld = B->ld;
  B->diag.alloc(ld);
  B->col.alloc(std::max<int64_t>((int64_t)B->K * ld, 1));
  B->val.alloc(std::max<int64_t>((int64_t)B->K * ld, 1) * B->vs);
  NC_CHECK_HIP(hipMemsetAsync(B->diag.p, 0, sizeof(double) * ld, s));
  NC_CHECK_HIP(hipMemsetAsync(B->col.p, 0, sizeof(int32_t) * B->K * ld, s));
  NC_CHECK_HIP(hipMemsetAsync(B->val.p, 0, B->vs * B->K * ld, s));
  B->gran_inst.alloc(B->grans);
  B->d_row_off.alloc(n);
  B->d_nstates.alloc(n);
  B->d_nrows.alloc(n);
  B->d_gran_start.alloc(n + 1);
  NC_CHECK_HIP(hipMemcpyAsync(B->gran_inst.p, gran_inst.data(), sizeof(int32_t) * B->grans, hipMemcpyHostToDevice, s));
  NC_CHECK_HIP(hipMemcpyAsync(B->d_row_off.p, B->row_off.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, s));
  NC_CHECK_HIP(hipMemcpyAsync(B->d_nstates.p, B->nstates.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, s));
  NC_CHECK_HIP(hipMemcpyAsync(B->d_nrows.p, B->nrows.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, s));
  NC_CHECK_HIP(hipMemcpyAsync(B->d_gran_start.p, B->gran_start.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, s));
  if (B->precision == NC_FP32)
    batch_build<float>(B.get(), mats);
  else
    batch_build<double>(B.get(), mats);
  // integrator buffers
  B->u.alloc(ld);
  B->ytmp.alloc(ld);
  B->k.alloc((size_t)7 * ld);
  B->vin.alloc(ld * B->vs);
  B->partials.alloc(B->grans);
  B->err.alloc(n);
  NC_CHECK_HIP(hipMemsetAsync(B->u.p, 0, sizeof(double) * ld, s));
  NC_CHECK_HIP(hipMemsetAsync(B->ytmp.p, 0, sizeof(double) * ld, s));
  NC_CHECK_HIP(hipMemsetAsync(B->k.p, 0, sizeof(double) * 7 * ld, s));
  NC_CHECK_HIP(hipMemsetAsync(B->vin.p, 0, B->vs * ld, s));
  NC_CHECK_HIP(hipMemsetAsync(B->partials.p, 0, sizeof(double) * B->grans, s));
  NC_CHECK_HIP(hipMemsetAsync(B->err.p, 0, sizeof(double) * n, s));
  B->c_t.alloc(n);
  B->c_hc.alloc(n);
  B->d_h.alloc(n);
  B->c_next.alloc(n);
  B->c_done.alloc(n);
  B->c_clamped.alloc(n);
  B->d_accept.alloc(n);
  B->d_snap.alloc(n);
  B->c_status.alloc(n);
  B->c_steps.alloc(n);
  B->c_rej.alloc(n);
  B->c_off.alloc(n + 1);
  B->n_agran.alloc(1);
  B->counters.alloc(2);
  B->agran.alloc(B->grans);
  B->par.alloc(1);
  NC_CHECK_HIP(hipMemsetAsync(B->par.p, 0, sizeof(Params), s));
  NC_CHECK_HIP(hipMemsetAsync(B->d_h.p, 0, sizeof(double) * n, s));
  NC_CHECK_HIP(hipMemsetAsync(B->c_clamped.p, 0, sizeof(int32_t) * n, s));
  NC_CHECK_HIP(hipMemsetAsync(B->n_agran.p, 0, sizeof(uint32_t), s));
  NC_CHECK_HIP(hipMemsetAsync(B->counters.p, 0, sizeof(uint32_t) * 2, s));
  B->cnt_pin = static_cast<uint32_t*>(pinned_alloc(sizeof(uint32_t) * 4));
  NC_CHECK_HIP(hipStreamSynchronize(s));
  return B.release();
}

Dev dev_view(nc_batch_s* B) {
  return Dev{B->gran_inst.p, B->d_row_off.p, B->d_nstates.p, B->d_h.p, B->d_accept.p, B->d_snap.p,
             B->agran.p,     B->n_agran.p};
}

Ctl ctl_view(nc_batch_s* B) {
  Ctl c{};
  c.n = B->n;
  c.par = B->par.p;
  c.t_out = B->t_out.p;
  c.err = B->err.p;
  c.nrows = B->d_nrows.p;
  c.gran_start = B->d_gran_start.p;
  c.t = B->c_t.p;
  c.hc = B->c_hc.p;
  c.htry = B->d_h.p;
  c.next = B->c_next.p;
  c.done = B->c_done.p;
  c.clamped = B->c_clamped.p;
  c.accept = B->d_accept.p;
  c.snap = B->d_snap.p;
  c.steps = B->c_steps.p;
  c.rejected = B->c_rej.p;
  c.status = B->c_status.p;
  c.off = B->c_off.p;
  c.agran = B->agran.p;
  c.n_agran = B->n_agran.p;
  c.counters = B->counters.p;
  return c;
}

// blocks of a list-driven launch: every unit of R granules once, at most kMaxRowBlocks
unsigned row_grid(const nc_batch_s* B, int R) {
  return (unsigned)std::min<int64_t>((B->grans + R - 1) / R, kMaxRowBlocks);
}

template <typename VT, int R>
void launch_matvec(nc_batch_s* B, const VT* v, double* out) {
  const Dev d = dev_view(B);
  hipStream_t s = B->stream.s;
  const SinkSegs sg{(int32_t)B->nseg, B->seg_inst.p, B->seg_row.p, B->seg_off.p, B->sink_src.p};
  const unsigned grid = row_grid(B, R) + (unsigned)B->nseg;
  auto go = [&](auto kern) {
    kern<<<grid, kBlock, 0, s>>>(d, B->ld, B->K, B->diag.p, B->col.p, reinterpret_cast<const VT*>(B->val.p), v, out,
                                 sg, reinterpret_cast<const VT*>(B->sink_val.p), B->dia_off.p, B->dia_inst.p, B->K);
  };
  auto by_k = [&](auto dm) {
    constexpr int DM = decltype(dm)::value;
    switch (B->K) {
      case 1: go(k_bmatvec<VT, R, 1, DM>); break;
      case 2: go(k_bmatvec<VT, R, 2, DM>); break;
      case 3: go(k_bmatvec<VT, R, 3, DM>); break;
      case 4: go(k_bmatvec<VT, R, 4, DM>); break;
      case 5: go(k_bmatvec<VT, R, 5, DM>); break;
      case 6: go(k_bmatvec<VT, R, 6, DM>); break;
      case 7: go(k_bmatvec<VT, R, 7, DM>); break;
      case 8: go(k_bmatvec<VT, R, 8, DM>); break;
      default: go(k_bmatvec<VT, R, 0, 0>); break;
    }
  };
  if (B->any_dia)
    by_k(std::integral_constant<int, 1>{});
  else
    by_k(std::integral_constant<int, 0>{});
  NC_CHECK_HIP(hipGetLastError());
}

template <typename VT, int R>
void launch_stage(nc_batch_s* B, int stage, unsigned g) {
  const Dev d = dev_view(B);
  hipStream_t s = B->stream.s;
  VT* vin = reinterpret_cast<VT*>(B->vin.p);
  auto go = [&](auto kern) { kern<<<g, kBlock, 0, s>>>(d, B->ld, B->u.p, B->k.p, B->ytmp.p, vin); };
  switch (stage) {
    case 0: go(k_bstage<VT, R, 0>); break;
    case 1: go(k_bstage<VT, R, 1>); break;
    case 2: go(k_bstage<VT, R, 2>); break;
    case 3: go(k_bstage<VT, R, 3>); break;
    case 4: go(k_bstage<VT, R, 4>); break;
    case 5: go(k_bstage<VT, R, 5>); break;
    case 6: go(k_bstage<VT, R, 6>); break;
    default: throw Error(NC_ERR_INTERNAL, "bad RK stage");
  }
}

void launch_control(nc_batch_s* B, int init) {
  k_bcontrol<<<1, kCtlBlock, 0, B->stream.s>>>(ctl_view(B), init);
  NC_CHECK_HIP(hipGetLastError());
}

// accept/snapshot of the previous step (fused with stage 1), then one full DP5 attempt,
// the error fold and (unless frozen) the controller
template <typename VT, int R>
void iteration_r(nc_batch_s* B, bool control) {
  const Dev d = dev_view(B);
  hipStream_t s = B->stream.s;
  const unsigned g = row_grid(B, R);
  k_baccept_stage1<VT, R><<<g, kBlock, 0, s>>>(d, B->ld, B->u.p, B->ytmp.p, B->k.p, B->snaps.p,
                                               reinterpret_cast<VT*>(B->vin.p));
  for (int st = 1; st < 7; ++st) {
    if (st > 1) launch_stage<VT, R>(B, st, g);
    launch_matvec<VT, R>(B, reinterpret_cast<const VT*>(B->vin.p), B->k.p + (int64_t)st * B->ld);
  }
  k_berr<R><<<g, kBlock, 0, s>>>(d, B->ld, B->par.p, B->u.p, B->ytmp.p, B->k.p, B->partials.p);
  k_bfold<<<(unsigned)((B->n + kBlock / 64 - 1) / (kBlock / 64)), kBlock, 0, s>>>(B->n, B->d_gran_start.p, B->d_h.p,
                                                                                B->partials.p, B->err.p);
  NC_CHECK_HIP(hipGetLastError());
  if (control) launch_control(B, 0);
}

template <typename VT>
void iteration(nc_batch_s* B, bool control) {
  switch (B->rows_per_thread) {
    case 1: iteration_r<VT, 1>(B, control); break;
    case 2: iteration_r<VT, 2>(B, control); break;
    default: iteration_r<VT, 4>(B, control); break;
  }
}

// k0 = A u for every listed instance (stage 0 copies u into the mat-vec input)
template <typename VT, int R>
void first_stage_r(nc_batch_s* B) {
  launch_stage<VT, R>(B, 0, row_grid(B, R));
  launch_matvec<VT, R>(B, reinterpret_cast<const VT*>(B->vin.p), B->k.p);
}

template <typename VT>
void first_stage(nc_batch_s* B) {
  switch (B->rows_per_thread) {
    case 1: first_stage_r<VT, 1>(B); break;
    case 2: first_stage_r<VT, 2>(B); break;
    default: first_stage_r<VT, 4>(B); break;
  }
}

void flush_accept(nc_batch_s* B) {
  hipStream_t s = B->stream.s;
  const int R = B->rows_per_thread;
  const Dev d = dev_view(B);
  const unsigned g = row_grid(B, R);
  if (R == 1)
    k_baccept<1><<<g, kBlock, 0, s>>>(d, B->ld, B->u.p, B->ytmp.p, B->k.p, B->snaps.p);
  else if (R == 2)
    k_baccept<2><<<g, kBlock, 0, s>>>(d, B->ld, B->u.p, B->ytmp.p, B->k.p, B->snaps.p);
  else
    k_baccept<4><<<g, kBlock, 0, s>>>(d, B->ld, B->u.p, B->ytmp.p, B->k.p, B->snaps.p);
  NC_CHECK_HIP(hipGetLastError());
}

void drop_graph(nc_batch_s* B) {
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  B->exec = nullptr;
  B->graph = nullptr;
}

// kChunk iterations, each ending with the controller
template <typename VT>
void run_chunk(nc_batch_s* B, bool graphs) {
  hipStream_t s = B->stream.s;
  if (!graphs) {
    for (int c = 0; c < kChunk; ++c) iteration<VT>(B, true);
    return;
  }
  if (B->exec && B->graph_R != B->rows_per_thread) drop_graph(B);
  if (!B->exec) {
    NC_CHECK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      for (int c = 0; c < kChunk; ++c) iteration<VT>(B, true);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    NC_CHECK_HIP(hipStreamEndCapture(s, &B->graph));
    NC_CHECK_HIP(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
    B->graph_R = B->rows_per_thread;
  }
  NC_CHECK_HIP(hipGraphLaunch(B->exec, s));
}

template <typename VT>
void batch_solve_t(nc_batch_s* B, const nc_solve_opts& o, const double* u0, int n_out, const double* t_out,
                   nc_solve_stats* st) {
  hipStream_t s = B->stream.s;
  const int n = B->n;
  const bool graphs = o.no_graphs == 0;
  Params* hp = static_cast<Params*>(host_stage(1));
  *hp = Params{o.atol > 0 ? o.atol : 1e-14, o.rtol > 0 ? o.rtol : 1e-8, o.h0, o.max_steps, n_out};
  NC_CHECK_HIP(hipMemcpyAsync(B->par.p, hp, sizeof(Params), hipMemcpyHostToDevice, s));
  // initial state, output times and snapshot storage (the captured chunk holds the
  // snapshot and output-time pointers: drop it with the old buffers)
  NC_CHECK_HIP(hipMemsetAsync(B->u.p, 0, sizeof(double) * B->ld, s));
  int64_t hoff = 0;
  for (int i = 0; i < n; ++i) {
    NC_CHECK_HIP(hipMemcpyAsync(B->u.p + B->row_off[i], u0 + hoff, sizeof(double) * B->nrows[i], hipMemcpyHostToDevice, s));
    hoff += B->nrows[i];
  }
  if ((size_t)std::max(n_out, 1) * B->ld > B->snaps.n || n_out > B->n_out_cap) {
    drop_graph(B);
    B->snaps.reserve((size_t)std::max(n_out, 1) * B->ld);
    B->t_out.reserve(std::max(n_out, 1));
    B->n_out_cap = std::max(n_out, 1);
  }
  B->n_out = n_out;
  if (n_out) NC_CHECK_HIP(hipMemcpyAsync(B->t_out.p, t_out, sizeof(double) * n_out, hipMemcpyHostToDevice, s));
  NC_CHECK_HIP(hipMemsetAsync(B->counters.p, 0, sizeof(uint32_t) * 2, s));
  launch_control(B, 1);  // initial controller state, first steps and list
  first_stage<VT>(B);
  NC_CHECK_HIP(hipMemcpyAsync(B->cnt_pin, B->counters.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  NC_CHECK_HIP(hipStreamSynchronize(s));
  int64_t chunks = 0, host_syncs = 1;
  // every instance ends by its own limits (status 2 or 3) first; this bound only keeps a
  // pathological controller from queueing work forever (8M iterations without max_steps)
  const int64_t max_chunks = o.max_steps > 0 ? o.max_steps / kChunk + 2 : (int64_t)1 << 20;
  if (B->cnt_pin[0] != 0) {
    // two chunks in flight: while one runs, the host reads the counter of the one before
    hipEvent_t ev[2];
    NC_CHECK_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    NC_CHECK_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    try {
      for (;;) {
        const int slot = (int)(chunks & 1);
        run_chunk<VT>(B, graphs);
        NC_CHECK_HIP(hipMemcpyAsync(B->cnt_pin + slot, B->counters.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        NC_CHECK_HIP(hipEventRecord(ev[slot], s));
        ++chunks;
        if (chunks >= 2) {
          NC_CHECK_HIP(hipEventSynchronize(ev[slot ^ 1]));
          ++host_syncs;
          if (B->cnt_pin[slot ^ 1] == 0) break;
        }
        // a bound every instance's own limits reach first (status 2 or 3 ends an instance)
        NC_REQUIRE(chunks < max_chunks, NC_ERR_NUMERIC, "batched RK exceeded max_steps");
      }
    } catch (...) {
      (void)hipStreamSynchronize(s);
      (void)hipEventDestroy(ev[0]);
      (void)hipEventDestroy(ev[1]);
      throw;
    }
    NC_CHECK_HIP(hipStreamSynchronize(s));
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
  }
  // apply the accepts and snapshots the last controller asked for
  flush_accept(B);
  std::vector<int32_t> status(n);
  std::vector<int64_t> steps(n), rej(n);
  NC_CHECK_HIP(hipMemcpyAsync(status.data(), B->c_status.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
  NC_CHECK_HIP(hipMemcpyAsync(steps.data(), B->c_steps.p, sizeof(int64_t) * n, hipMemcpyDeviceToHost, s));
  NC_CHECK_HIP(hipMemcpyAsync(rej.data(), B->c_rej.p, sizeof(int64_t) * n, hipMemcpyDeviceToHost, s));
  NC_CHECK_HIP(hipMemcpyAsync(B->cnt_pin + 2, B->counters.p + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  NC_CHECK_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < n; ++i) {
    if (status[i] == 1) throw Error(NC_ERR_NUMERIC, "non-finite error estimate in batched instance " + std::to_string(i));
    if (status[i] == 2) throw Error(NC_ERR_NUMERIC, "batched RK step size underflow in instance " + std::to_string(i));
    if (status[i] == 3) throw Error(NC_ERR_NUMERIC, "batched RK exceeded max_steps in instance " + std::to_string(i));
  }
  if (st) {
    *st = nc_solve_stats{};
    for (int i = 0; i < n; ++i) st->steps += steps[i], st->rejected += rej[i];
    st->host_syncs = host_syncs;
    st->matvecs = (int64_t)B->cnt_pin[2] * 6;  // batched mat-vec launches with work
  }
}

}  // namespace

extern "C" {

int nc_batch_create_models(int32_t n, const nc_model* models, const int32_t* const* states, const int64_t* n_states,
                           int32_t precision, int32_t n_workers, nc_batch* out) {
  return guard([&] {
    NC_REQUIRE(out && models && states && n_states && n >= 1, NC_ERR_ARG, "bad arguments");
    NC_REQUIRE(precision == NC_FP32 || precision == NC_FP64, NC_ERR_ARG, "precision must be NC_FP32 or NC_FP64");
    for (int32_t i = 0; i < n; ++i) {
      NC_REQUIRE(models[i], NC_ERR_ARG, "null model of instance " + std::to_string(i));
      NC_REQUIRE(n_states[i] >= 1 && states[i], NC_ERR_ARG, "empty state list of instance " + std::to_string(i));
    }
    int dev = 0;
    NC_CHECK_HIP(hipGetDevice(&dev));
    const int W = std::max(1, std::min<int>(n, n_workers > 0 ? n_workers : kBuildWorkers));
    std::vector<std::unique_ptr<nc_space_s>> spaces(n);
    std::vector<std::unique_ptr<nc_matrix_s>> mats(n);
    std::atomic<int32_t> next{0};
    std::atomic<bool> failed{false};
    std::mutex err_mu;
    int err_code = NC_OK;
    std::string err_msg;
    auto fail = [&](int32_t i, int code, const char* what) {
      std::lock_guard<std::mutex> lk(err_mu);
      if (err_code == NC_OK) {
        err_code = code;
        err_msg = i >= 0 ? "instance " + std::to_string(i) + ": " + what : std::string(what);
      }
      failed = true;
    };
    // each worker builds whole instances (state space, then generator) on streams of its own
    auto work = [&] {
      if (hipSetDevice(dev) != hipSuccess) {
        fail(-1, NC_ERR_HIP, "hipSetDevice failed on a worker thread");
        return;
      }
      PrivateStreams scope;
      for (;;) {
        const int32_t i = next.fetch_add(1);
        if (i >= n || failed) break;
        try {
          nc_model_s* m = models[i];
          spaces[i].reset(space_create(m->ns, m->nr, m->stoich.data(), n_states[i], states[i]));
          mats[i].reset(matrix_create(spaces[i].get(), m, precision));
          NC_CHECK_HIP(hipStreamSynchronize(spaces[i]->stream.s));
          NC_CHECK_HIP(hipStreamSynchronize(mats[i]->stream.s));
          // this instance is complete and only read by the batch's copies from here on:
          // its streams serve the next instance (a worker keeps a few streams in use, not
          // two per instance), and later releases need not wait for it
          spaces[i]->stream.release_now();
          mats[i]->stream.release_now();
          streams_settled();
        } catch (const Error& e) {
          fail(i, e.code, e.what());
        } catch (const std::bad_alloc&) {
          fail(i, NC_ERR_ALLOC, "host allocation failed");
        } catch (const std::exception& e) {
          fail(i, NC_ERR_INTERNAL, e.what());
        }
      }
    };
    std::vector<std::thread> pool;
    pool.reserve(W - 1);
    try {
      for (int w = 1; w < W; ++w) pool.emplace_back(work);
    } catch (const std::system_error&) {
      // fewer threads than asked for: the ones started (and this one) take every instance
    }
    work();
    for (auto& th : pool) th.join();
    if (failed) {
      NC_CHECK_HIP(hipDeviceSynchronize());
      throw Error(err_code, err_msg);
    }
    std::vector<nc_matrix> raw(n);
    for (int32_t i = 0; i < n; ++i) raw[i] = mats[i].get();
    nc_batch_s* B = batch_create(n, raw.data());
    // the batch has copied the generators (on its own stream): the instances go once the
    // copies are done
    const hipError_t e = hipStreamSynchronize(B->stream.s);
    if (e != hipSuccess) {
      delete B;
      throw Error(NC_ERR_HIP, std::string("batch copy failed: ") + hipGetErrorString(e));
    }
    {
      // every instance's work is complete (synchronised above), so their buffers go back
      // to the cache without a device-wide synchronisation each
      PrivateStreams scope;
      mats.clear();
      spaces.clear();
    }
    *out = B;
  });
}

int nc_batch_create(int32_t n, const nc_matrix* mats, nc_batch* out) {
  return guard([&] {
    NC_REQUIRE(out, NC_ERR_ARG, "null argument");
    *out = batch_create(n, mats);
  });
}

int nc_batch_destroy(nc_batch B) {
  return guard([&] { delete B; });
}

int nc_batch_autotune(nc_batch B, int32_t reps, int32_t* rows_per_thread, double* best_ms) {
  return guard([&] {
    NC_REQUIRE(B && reps >= 1, NC_ERR_ARG, "bad arguments");
    hipStream_t s = B->stream.s;
    // every instance listed with a tiny step and the controller left out: nothing is
    // accepted or saved, and every iteration does the full batch's work
    Params* hp = static_cast<Params*>(host_stage(1));
    *hp = Params{1e-14, 1e-8, 0.0, 0, 0};
    NC_CHECK_HIP(hipMemcpyAsync(B->par.p, hp, sizeof(Params), hipMemcpyHostToDevice, s));
    k_bfreeze<<<grid_for(std::max<int64_t>(B->grans, B->n), 256), 256, 0, s>>>(
        B->n, 1e-9, B->d_gran_start.p, B->d_h.p, B->d_accept.p, B->d_snap.p, B->agran.p, B->n_agran.p);
    NC_CHECK_HIP(hipGetLastError());
    if (!B->snaps.p) B->snaps.reserve(B->ld);  // the accept kernel takes a snapshot pointer
    hipEvent_t e0, e1;
    NC_CHECK_HIP(hipEventCreate(&e0));
    NC_CHECK_HIP(hipEventCreate(&e1));
    double best = 1e300;
    int best_r = B->rows_per_thread;
    for (int r : {1, 2, 4}) {
      B->rows_per_thread = r;
      auto it = [&] {
        if (B->precision == NC_FP32)
          iteration<float>(B, false);
        else
          iteration<double>(B, false);
      };
      it();  // warm-up
      NC_CHECK_HIP(hipEventRecord(e0, s));
      for (int k = 0; k < reps; ++k) it();
      NC_CHECK_HIP(hipEventRecord(e1, s));
      NC_CHECK_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      NC_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms / reps < best) best = ms / reps, best_r = r;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    B->rows_per_thread = best_r;
    drop_graph(B);  // a captured chunk was recorded with the previous launch shape
    if (rows_per_thread) *rows_per_thread = best_r;
    if (best_ms) *best_ms = best;
  });
}

int nc_batch_set_rows_per_thread(nc_batch B, int32_t rows) {
  return guard([&] {
    NC_REQUIRE(B, NC_ERR_ARG, "null batch");
    NC_REQUIRE(rows == 1 || rows == 2 || rows == 4, NC_ERR_ARG, "rows per thread must be 1, 2 or 4");
    if (rows != B->rows_per_thread) drop_graph(B);
    B->rows_per_thread = rows;
  });
}

int nc_batch_info(nc_batch B, int64_t* total_rows, int64_t* instance_rows) {
  return guard([&] {
    NC_REQUIRE(B, NC_ERR_ARG, "null batch");
    if (total_rows) *total_rows = B->ld;
    if (instance_rows)
      for (int i = 0; i < B->n; ++i) instance_rows[i] = B->nrows[i];
  });
}

int nc_batch_solve(nc_batch B, const nc_solve_opts* opts, const double* u0, int32_t n_out, const double* t_out,
                   nc_solve_stats* st) {
  return guard([&] {
    NC_REQUIRE(B && u0 && n_out >= 1 && t_out, NC_ERR_ARG, "bad arguments");
    for (int k = 1; k < n_out; ++k) NC_REQUIRE(t_out[k] > t_out[k - 1], NC_ERR_ARG, "output times must increase");
    nc_solve_opts o{};
    if (opts)
      o = *opts;
    else
      nc_solve_opts_default(&o);
    if (B->precision == NC_FP32)
      batch_solve_t<float>(B, o, u0, n_out, t_out, st);
    else
      batch_solve_t<double>(B, o, u0, n_out, t_out, st);
  });
}

int nc_batch_get_snapshots(nc_batch B, int32_t k, double* u) {
  return guard([&] {
    NC_REQUIRE(B && u && k >= 0 && k < B->n_out, NC_ERR_ARG, "index out of range");
    hipStream_t s = B->stream.s;
    std::vector<int64_t> off(B->n);
    int64_t total = 0;
    for (int i = 0; i < B->n; ++i) off[i] = total, total += B->nrows[i];
    DevBuf<int64_t> d_off(B->n);
    DevBuf<double> packed(total);
    NC_CHECK_HIP(hipMemcpyAsync(d_off.p, off.data(), sizeof(int64_t) * B->n, hipMemcpyHostToDevice, s));
    k_bpack<<<B->n, kBlock, 0, s>>>(B->snaps.p + (size_t)k * B->ld, B->d_row_off.p, B->d_nrows.p, d_off.p, packed.p);
    NC_CHECK_HIP(hipGetLastError());
    copy_d2h(u, packed.p, sizeof(double) * total, s);  // ends with the stream synchronised
  });
}

int nc_batch_get_snapshot(nc_batch B, int32_t inst, int32_t k, double* u) {
  return guard([&] {
    NC_REQUIRE(B && u && inst >= 0 && inst < B->n && k >= 0 && k < B->n_out, NC_ERR_ARG, "index out of range");
    NC_CHECK_HIP(hipMemcpy(u, B->snaps.p + (size_t)k * B->ld + B->row_off[inst], sizeof(double) * B->nrows[inst],
                           hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
