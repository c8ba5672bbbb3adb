// Forward-sensitivity FSP operator: ForwardSensFspMatrixSparse and its matvec!
// (src/forwardsensfspmatrix/forwardsensfspmatrixsparse/sensfspmatrixsparse.jl:31-142).
//
// One fused kernel applies A to all P+1 blocks while streaming A's slots once, and adds
// the derivative terms dA_p v_0 from per-pair arrays. The matrix-side bytes of a launch
// are those of one mat-vec; the vector-side bytes grow with the block count.
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

#include "nc_sens.h"

nc_sensmatrix_s::~nc_sensmatrix_s() { delete A; }

namespace nc {
namespace {

constexpr int kBlock = 256;

struct PairMeta {
  int32_t r, arr, cslot, q;
  int64_t sink_off;
};

// One derivative term of the row part: the pair's state-factor derivative (kind 0:
// ddiag/dpred row idx, predecessor column slot cslot, scaled by dscale[pair]) or its
// time-factor derivative (kind 1: time-varying slot idx, scaled by sfscale[pair]), added
// to accumulator kk of the wave that owns the pair's block (kind -1: padding).
struct Term {
  int8_t kk, pad;
  int16_t kind;
  int16_t idx, cslot;
  int16_t pair, pad2;
};
constexpr int kMaxTerms = 2 * kMaxPairs;

template <typename VT>
struct SensArgs {
  int64_t n, ld, vstride, ostride;
  int32_t K, ntv, nr, nb;
  int32_t fold;  // tile fold (1: none)
  const double* diag;
  const int32_t* col;
  const VT* val;
  const VT* tvd;
  const int32_t* tvc;
  const VT* tvv;
  const int64_t* sink_off;
  const int32_t* sink_src;
  const VT* sink_val;
  const VT* ddiag;
  const VT* dpred;
  const int32_t* dcol;
  const VT* dsink;
  const PairMeta* meta;  // pairs sorted by parameter
  const Term* terms;     // derivative terms of the row part, grouped by wave
  int32_t tstart[kBlock / 64 + 1];  // terms of wave w: tstart[w] .. tstart[w+1)
  int32_t bstart[kBlock / 64 + 1];  // blocks of wave w: bstart[w] .. bstart[w+1)
  int32_t pstart[kMaxSensBlocks + 1];  // pairs of block b (b >= 1): pstart[b] .. pstart[b+1)
  uint32_t dia_mask;                   // bit k: slot k's column is i + dia_off[k]
  int32_t dia_off[kMaxReactions];
  double tv_scale[kMaxReactions];
  double sink_scale[kMaxReactions];
  double dscale[kMaxPairs];   // multiplies the pair's state-factor derivative term
  double sfscale[kMaxPairs];  // multiplies the time-factor derivative term (separable)
  // DP5(4) epilogue (EPI kernels): the next stage input of row i of block b, index
  // b * rk_stride + i of the stage vectors, u, ytmp and vin
  const double* rk_k[6];  // k_j (a duplicate of rk_k[0] where the mask bit is clear)
  double rk_a[7];
  double rk_h;
  const double* rk_u;
  double* rk_ytmp;        // the last stage: y_{n+1} in FP64 too (else null)
  VT* rk_vin;
  int64_t rk_stride;
  int32_t rk_mask;
};

// The DP5(4) stage combination of one (block, row) entry at index e of the stage vectors,
// from the value knew the mat-vec has just formed: the fused multiply-adds of the separate
// stage pass, in its order (the k loads are unconditional; a clear mask bit points at k_0
// and its term is skipped by a select).
template <typename VT>
__device__ inline void sens_rk_epilogue(const SensArgs<VT>& a, int64_t e, double knew) {
  double kv[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) kv[j] = a.rk_k[j][e];
  const double uv = a.rk_u[e];
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const double t = __fma_rn(a.rk_a[j], kv[j], acc);
    acc = ((a.rk_mask >> j) & 1) ? t : acc;
  }
  acc = __fma_rn(a.rk_a[6], knew, acc);
  const double y = __fma_rn(a.rk_h, acc, uv);
  a.rk_vin[e] = (VT)y;
  if (a.rk_ytmp) a.rk_ytmp[e] = y;
}

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// The row part of one wave: blocks b0 .. b0 + M - 1 of the 64 rows i of its tile. Every
// load count is exact (no padded slot, block or term): the kernel is bound by how many
// loads it keeps in flight, so a padded load costs as much as a useful one. The loads of a
// row go out in batches whose addresses do not depend on each other (the slot columns and
// values with the row's own entries, then every gather for them; the time-varying slots two
// at a time; the derivative terms up to four at a time).
template <typename VT, int M>
__device__ __forceinline__ void sens_tv(const SensArgs<VT>& a, const VT* const (&vb)[M], const double (&vi)[M],
                                        double (&acc)[M], int q, int nq, int64_t i) {
  const int64_t ld = a.ld;
  if (nq == 2) {
    int32_t c[2];
    double d[2], w[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      c[u] = a.tvc[(q + u) * ld + i];
      d[u] = (double)a.tvd[(q + u) * ld + i];
      w[u] = (double)a.tvv[(q + u) * ld + i];
    }
    double g[2][M];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int kk = 0; kk < M; ++kk) g[u][kk] = (double)vb[kk][c[u]];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int kk = 0; kk < M; ++kk) acc[kk] += a.tv_scale[q + u] * (d[u] * vi[kk] + w[u] * g[u][kk]);
  } else {
    const int32_t c = a.tvc[q * ld + i];
    const double d = (double)a.tvd[q * ld + i], w = (double)a.tvv[q * ld + i];
    double g[M];
#pragma unroll
    for (int kk = 0; kk < M; ++kk) g[kk] = (double)vb[kk][c];
#pragma unroll
    for (int kk = 0; kk < M; ++kk) acc[kk] += a.tv_scale[q] * (d * vi[kk] + w * g[kk]);
  }
}

// NT derivative terms dA_p v_0 (pair p's block is the wave's accumulator T.kk)
template <typename VT, int M, int NT>
__device__ __forceinline__ void sens_terms(const SensArgs<VT>& a, const Term* s_terms, const VT* __restrict__ v,
                                           double v0, double (&acc)[M], int t0, int64_t i) {
  const int64_t ld = a.ld;
  int32_t c[NT], kk_of[NT];
  double d[NT], w[NT], g[NT], sc[NT];
#pragma unroll
  for (int u = 0; u < NT; ++u) {
    const Term T = s_terms[t0 + u];
    const bool st = T.kind == 0;
    sc[u] = st ? a.dscale[T.pair] : a.sfscale[T.pair];
    kk_of[u] = T.kk;
    const int32_t* cp = st ? a.dcol + (int64_t)T.cslot * ld : a.tvc + (int64_t)T.idx * ld;
    const VT* dp = st ? a.ddiag + (int64_t)T.idx * ld : a.tvd + (int64_t)T.idx * ld;
    const VT* wp = st ? a.dpred + (int64_t)T.idx * ld : a.tvv + (int64_t)T.idx * ld;
    c[u] = cp[i];
    d[u] = (double)dp[i];
    w[u] = (double)wp[i];
  }
#pragma unroll
  for (int u = 0; u < NT; ++u) g[u] = (double)v[c[u]];
#pragma unroll
  for (int u = 0; u < NT; ++u) {
    const double add = sc[u] * (d[u] * v0 + w[u] * g[u]);
#pragma unroll
    for (int kk = 0; kk < M; ++kk) acc[kk] += kk == kk_of[u] ? add : 0.0;
  }
}

template <typename VT, int M, int K, bool EPI>
__device__ __forceinline__ void sens_rows(const SensArgs<VT>& a, const Term* s_terms, const VT* __restrict__ v,
                                          double* __restrict__ out, int wave, int64_t i) {
  const int64_t ld = a.ld, vs = a.vstride;
  const int b0 = a.bstart[wave];
  const VT* vb[M];
#pragma unroll
  for (int k = 0; k < M; ++k) vb[k] = v + (int64_t)(b0 + k) * vs;
  double acc[M], vi[M];
  const double dg = a.diag[i];
  const double v0 = (double)v[i];
#pragma unroll
  for (int k = 0; k < M; ++k) vi[k] = (double)vb[k][i];
  // the slot's column: computed (i + dia_off, clamped: rows whose value is 0 may point
  // anywhere) or loaded
  auto col = [&](int k) -> int32_t {
    if ((a.dia_mask >> k) & 1u) {
      const int64_t j = i + a.dia_off[k];
      return (int32_t)(j < 0 ? 0 : (j >= a.n ? a.n - 1 : j));
    }
    return a.col[k * ld + i];
  };
  if constexpr (K > 0) {
    int32_t c[K];
    double w[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      c[k] = col(k);
      w[k] = (double)a.val[k * ld + i];
    }
#pragma unroll
    for (int kk = 0; kk < M; ++kk) acc[kk] = dg * vi[kk];
#pragma unroll
    for (int k = 0; k < K; ++k) {
#pragma unroll
      for (int kk = 0; kk < M; ++kk) acc[kk] += w[k] * (double)vb[kk][c[k]];
    }
  } else {
#pragma unroll
    for (int kk = 0; kk < M; ++kk) acc[kk] = dg * vi[kk];
    for (int k = 0; k < a.K; ++k) {
      const int32_t c = col(k);
      const double w = (double)a.val[k * ld + i];
#pragma unroll
      for (int kk = 0; kk < M; ++kk) acc[kk] += w * (double)vb[kk][c];
    }
  }
  int q = 0;
  for (; q + 2 <= a.ntv; q += 2) sens_tv<VT, M>(a, vb, vi, acc, q, 2, i);
  if (q < a.ntv) sens_tv<VT, M>(a, vb, vi, acc, q, 1, i);
  int t = a.tstart[wave];
  const int t1 = a.tstart[wave + 1];
  for (; t + 4 <= t1; t += 4) sens_terms<VT, M, 4>(a, s_terms, v, v0, acc, t, i);
  switch (t1 - t) {
    case 3: sens_terms<VT, M, 3>(a, s_terms, v, v0, acc, t, i); break;
    case 2: sens_terms<VT, M, 2>(a, s_terms, v, v0, acc, t, i); break;
    case 1: sens_terms<VT, M, 1>(a, s_terms, v, v0, acc, t, i); break;
    default: break;
  }
#pragma unroll
  for (int kk = 0; kk < M; ++kk) out[(int64_t)(b0 + kk) * a.ostride + i] = acc[kk];
  if constexpr (EPI) {
#pragma unroll
    for (int kk = 0; kk < M; ++kk) sens_rk_epilogue<VT>(a, (int64_t)(b0 + kk) * a.rk_stride + i, acc[kk]);
  }
}

// One workgroup per tile of 64 rows; its four waves share the tile's rows and split the
// P + 1 vector blocks in contiguous chunks (MH = ceil((P + 1) / 4) blocks for the first
// waves, MH - 1 for the rest), so a thread holds at most 8 block accumulators instead of
// P + 1. A's slot arrays are read by the four waves of one compute unit (one HBM read, the
// rest from its caches). The derivative terms dA_p v_0 of a wave's blocks come from a
// host-built list (Term), staged in LDS. The nb x nr sink rows come first in the grid, one
// workgroup each, so that their longer reductions overlap the row tiles.
//
// Tile order: tiles go to the XCDs in contiguous ranges (blockIdx.x modulo 8 selects the
// XCD), and within a range a fold F > 1 interleaves the F segments of tiles / F tiles each,
// so that rows one segment apart (a slow species' neighbours, such as the gene states of a
// one-hot model) are processed by the same XCD at about the same time and their gathers
// hit its L2 instead of going back to memory.
template <typename VT, int MH, int K, bool EPI = false>
__global__ __launch_bounds__(kBlock) void k_sens_matvec(SensArgs<VT> a, const VT* __restrict__ v,
                                                        double* __restrict__ out) {
  constexpr int W = kBlock / 64;
  const int nb = a.nb;
  const int64_t nsink = (int64_t)nb * a.nr, nsink8 = (nsink + 7) & ~int64_t(7);
  if ((int64_t)blockIdx.x >= nsink8) {
    // the derivative terms, staged in LDS (uniform reads; no load waits on them later)
    __shared__ Term s_terms[kMaxTerms];
    const int nterms = a.tstart[W];
    for (int t = threadIdx.x; t < nterms; t += kBlock) s_terms[t] = a.terms[t];
    __syncthreads();
    const int64_t tiles = (a.n + 63) / 64;
    const int64_t b = (int64_t)blockIdx.x - nsink8, q8 = tiles >> 3, r8 = tiles & 7, x = b & 7;
    int64_t tile = x * q8 + (x < r8 ? x : r8) + (b >> 3);
    if (a.fold > 1) {
      const int64_t seg = tiles / a.fold;  // a.fold * seg <= tiles; the rest keep their place
      if (tile < a.fold * seg) tile = (tile % a.fold) * seg + tile / a.fold;
    }
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t i = tile * 64 + (threadIdx.x & 63);
    if (i >= a.n) return;
    const int nblk = a.bstart[wave + 1] - a.bstart[wave];
    if (nblk == MH)
      sens_rows<VT, MH, K, EPI>(a, s_terms, v, out, wave, i);
    else if constexpr (MH > 1)
      if (nblk == MH - 1) sens_rows<VT, MH - 1, K, EPI>(a, s_terms, v, out, wave, i);
  } else if ((int64_t)blockIdx.x < nsink) {
    // sink row r of block bb: one workgroup, deterministic tree
    const int sb = (int)blockIdx.x;
    const int bb = sb / a.nr, r = sb % a.nr;
    const int64_t e0 = a.sink_off[r], e1 = a.sink_off[r + 1];
    const VT* vb = v + bb * a.vstride;
    // four entries per iteration, their loads clamped and issued together, then their
    // gathers (the thread's entries keep the order of a plain loop)
    constexpr int U = 4;
    double acc = 0.0;
    for (int64_t eb = e0 + threadIdx.x; eb < e1; eb += (int64_t)kBlock * U) {
      int32_t src[U];
      double w[U], g[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t e = eb + (int64_t)u * kBlock, ec = e < e1 ? e : e1 - 1;
        src[u] = a.sink_src[ec];
        w[u] = (double)a.sink_val[ec];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) g[u] = (double)vb[src[u]];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (eb + (int64_t)u * kBlock < e1) acc += w[u] * g[u];
    }
    acc *= a.sink_scale[r];
    if (bb >= 1) {
      for (int p = a.pstart[bb]; p < a.pstart[bb + 1]; ++p) {
        const PairMeta m = a.meta[p];
        if (m.r != r) continue;
        double s1 = 0.0, s2 = 0.0;
        const VT* ds = m.arr >= 0 ? a.dsink + m.sink_off : a.sink_val + e0;  // (dsink read only if arr >= 0)
        for (int64_t eb = e0 + threadIdx.x; eb < e1; eb += (int64_t)kBlock * U) {
          int32_t src[U];
          double d[U], w[U], g[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int64_t e = eb + (int64_t)u * kBlock, ec = e < e1 ? e : e1 - 1;
            src[u] = a.sink_src[ec];
            d[u] = (double)ds[ec - e0];
            w[u] = (double)a.sink_val[ec];
          }
#pragma unroll
          for (int u = 0; u < U; ++u) g[u] = (double)v[src[u]];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            if (eb + (int64_t)u * kBlock >= e1) continue;
            if (m.arr >= 0) s1 += d[u] * g[u];
            if (m.q >= 0) s2 += w[u] * g[u];
          }
        }
        acc += a.dscale[p] * s1 + a.sfscale[p] * s2;
      }
    }
    __shared__ double part[kBlock / 64];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double sum = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) sum += part[w];
      out[bb * a.ostride + a.n + r] = sum;
      if constexpr (EPI) sens_rk_epilogue<VT>(a, bb * a.rk_stride + a.n + r, sum);
    }
  }
}

struct PairFill {
  const nc_instr* code;
  int64_t len;
  int32_t r, arr, cslot;
};

// Pair arrays at time t (t matters for joint reactions only).
template <typename VT>
__global__ void k_sens_fill(int64_t n, int64_t ld, int ns, int nr, const int32_t* states, const int32_t* state_conn,
                            const int32_t* stoich, const double* theta, PairFill pf, double t, VT* ddiag, VT* dpred,
                            int32_t* dcol, unsigned long long* dmax_bits) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= ld) return;
  VT* dd = ddiag + pf.arr * ld;
  VT* dp = dpred + pf.arr * ld;
  int32_t* dc = dcol ? dcol + pf.cslot * ld : nullptr;
  if (i >= n) {
    dd[i] = (VT)0;
    dp[i] = (VT)0;
    if (dc) dc[i] = 0;
    return;
  }
  int32_t x[kMaxSpecies], y[kMaxSpecies];
  for (int s = 0; s < ns; ++s) x[s] = states[i * ns + s];
  const double di = prog_eval(pf.code, pf.len, t, x, theta);
  dd[i] = -(VT)di;
  const int32_t j = state_conn[i * nr + pf.r];
  if (j >= 0) {
    for (int s = 0; s < ns; ++s) y[s] = x[s] - stoich[s + pf.r * ns];
    dp[i] = (VT)prog_eval(pf.code, pf.len, t, y, theta);
  } else {
    dp[i] = (VT)0;
  }
  if (dc) dc[i] = j >= 0 ? j : (int32_t)i;
  if (dmax_bits) atomicMax(dmax_bits, (unsigned long long)__double_as_longlong(fabs(di)));
}

template <typename VT>
__global__ void k_sens_sink_fill(int64_t e0, int64_t e1, int ns, const int32_t* states, const int32_t* sink_src,
                                 const double* theta, PairFill pf, double t, VT* dsink) {
  const int64_t e = e0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= e1) return;
  int32_t x[kMaxSpecies];
  const int64_t j = sink_src[e];
  for (int s = 0; s < ns; ++s) x[s] = states[j * ns + s];
  dsink[e - e0] = (VT)prog_eval(pf.code, pf.len, t, x, theta);
}

template <typename VT>
void fill_pairs(nc_sensmatrix_s* SA, const int32_t* states, const int32_t* state_conn, double t, bool joint_only,
                unsigned long long* dmax_bits) {
  nc_matrix_s* A = SA->A;
  nc_model_s* m = SA->model;
  hipStream_t s = A->stream.s;
  std::vector<bool> col_done(SA->n_cslots, joint_only);
  for (const SensPair& p : SA->pairs) {
    if (p.arr < 0) continue;
    if (joint_only && p.kind != NC_PROP_JOINT) continue;
    PairFill pf{m->d_dcode.p + m->doffs[2 * (p.r * m->np + p.p)], m->dxlen(p.r, p.p), p.r, p.arr, p.cslot};
    const bool write_col = !col_done[p.cslot];
    col_done[p.cslot] = true;
    if (A->ld)
      k_sens_fill<VT><<<grid_for(A->ld, kBlock), kBlock, 0, s>>>(
          A->n, A->ld, A->ns, A->nr, states, state_conn, m->d_stoich.p, m->d_params.p, pf, t,
          reinterpret_cast<VT*>(SA->ddiag.p), reinterpret_cast<VT*>(SA->dpred.p), write_col ? SA->dcol.p : nullptr,
          dmax_bits);
    const int64_t e0 = A->sink_off_h[p.r], e1 = A->sink_off_h[p.r + 1];
    if (e1 > e0)
      k_sens_sink_fill<VT><<<grid_for(e1 - e0, kBlock), kBlock, 0, s>>>(
          e0, e1, A->ns, states, A->sink_src.p, m->d_params.p, pf, t,
          reinterpret_cast<VT*>(SA->dsink.p) + p.sink_off);
    NC_CHECK_HIP(hipGetLastError());
  }
}

template <typename VT>
void launch_sens(nc_sensmatrix_s* SA, double t, const void* v, int64_t vstride, double* out, int64_t ostride,
                 hipStream_t s, hipEvent_t e0, hipEvent_t e1, const RkEpilogue* epi) {
  nc_matrix_s* A = SA->A;
  nc_model_s* m = SA->model;
  SensArgs<VT> a{};
  a.n = A->n;
  a.ld = A->ld;
  a.vstride = vstride;
  a.ostride = ostride;
  a.K = A->K;
  a.ntv = (int32_t)A->tv_r.size();
  a.nr = A->nr;
  a.nb = SA->P + 1;
  a.diag = A->diag.p;
  a.col = A->col.p;
  a.val = reinterpret_cast<const VT*>(A->val.p);
  a.tvd = reinterpret_cast<const VT*>(A->tvd.p);
  a.tvc = A->tvc.p;
  a.tvv = reinterpret_cast<const VT*>(A->tvv.p);
  a.sink_off = A->sink_off.p;
  a.sink_src = A->sink_src.p;
  a.sink_val = reinterpret_cast<const VT*>(A->sink_val.p);
  a.ddiag = reinterpret_cast<const VT*>(SA->ddiag.p);
  a.dpred = reinterpret_cast<const VT*>(SA->dpred.p);
  a.dcol = SA->dcol.p;
  a.dsink = reinterpret_cast<const VT*>(SA->dsink.p);
  a.meta = reinterpret_cast<const PairMeta*>(SA->meta_dev.p);
  a.terms = reinterpret_cast<const Term*>(SA->terms_dev.p);
  for (int w = 0; w <= kBlock / 64; ++w) {
    a.tstart[w] = SA->tstart[w];
    a.bstart[w] = SA->bstart[w];
  }
  a.fold = SA->fold;
  a.dia_mask = A->dia_mask;
  for (size_t k = 0; k < A->dia_off.size() && k < (size_t)kMaxReactions; ++k) a.dia_off[k] = A->dia_off[k];
  matrix_scales(A, t, a.sink_scale, a.tv_scale);
  for (int b = 0; b <= SA->P + 1; ++b) a.pstart[b] = SA->pstart[b];
  for (size_t k = 0; k < SA->pairs.size(); ++k) {
    const SensPair& p = SA->pairs[k];
    a.dscale[k] = p.kind == NC_PROP_SEPARABLE ? m->tfactor(p.r, t) : 1.0;
    a.sfscale[k] = (p.q >= 0 && m->dtlen(p.r, p.p) > 0)
                       ? prog_eval(m->dtprog(p.r, p.p), m->dtlen(p.r, p.p), t, nullptr, m->params.data())
                       : 0.0;
  }
  const int64_t row_tiles = (A->n + 63) / 64;
  const unsigned grid = (unsigned)(row_tiles + (((int64_t)a.nb * A->nr + 7) & ~int64_t(7)));
  const VT* vv = reinterpret_cast<const VT*>(v);
  auto go = [&](auto kern) {
    if (e0)
      hipExtLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, s, e0, e1, 0, a, vv, out);
    else
      kern<<<grid, kBlock, 0, s>>>(a, vv, out);
  };
  static_assert(kMaxSensBlocks <= 8 * (kBlock / 64), "blocks per wave");
  if (epi) {
    // the stage epilogue kernels (FP32 storage, at most 4 blocks per wave: P + 1 <= 16)
    if (!sens_rk_fusable(SA) || !std::is_same<VT, float>::value)
      throw Error(NC_ERR_INTERNAL, "sensitivity mat-vec epilogue not available");
    if constexpr (std::is_same<VT, float>::value) {
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
      a.rk_stride = vstride;
      auto by_k_epi = [&](auto mh) {
        constexpr int M = decltype(mh)::value;
        switch (a.K) {
          case 1: go(k_sens_matvec<VT, M, 1, true>); break;
          case 2: go(k_sens_matvec<VT, M, 2, true>); break;
          case 3: go(k_sens_matvec<VT, M, 3, true>); break;
          case 4: go(k_sens_matvec<VT, M, 4, true>); break;
          case 5: go(k_sens_matvec<VT, M, 5, true>); break;
          case 6: go(k_sens_matvec<VT, M, 6, true>); break;
          case 7: go(k_sens_matvec<VT, M, 7, true>); break;
          case 8: go(k_sens_matvec<VT, M, 8, true>); break;
          default: go(k_sens_matvec<VT, M, 0, true>); break;
        }
      };
      switch ((a.nb + 3) / 4) {
        case 1: by_k_epi(std::integral_constant<int, 1>{}); break;
        case 2: by_k_epi(std::integral_constant<int, 2>{}); break;
        case 3: by_k_epi(std::integral_constant<int, 3>{}); break;
        default: by_k_epi(std::integral_constant<int, 4>{}); break;
      }
      NC_CHECK_HIP(hipGetLastError());
    }
    return;
  }
  auto by_k = [&](auto mh) {
    constexpr int M = decltype(mh)::value;
    switch (a.K) {
      case 1: go(k_sens_matvec<VT, M, 1>); break;
      case 2: go(k_sens_matvec<VT, M, 2>); break;
      case 3: go(k_sens_matvec<VT, M, 3>); break;
      case 4: go(k_sens_matvec<VT, M, 4>); break;
      case 5: go(k_sens_matvec<VT, M, 5>); break;
      case 6: go(k_sens_matvec<VT, M, 6>); break;
      case 7: go(k_sens_matvec<VT, M, 7>); break;
      case 8: go(k_sens_matvec<VT, M, 8>); break;
      default: go(k_sens_matvec<VT, M, 0>); break;
    }
  };
  // the most blocks a wave owns (the dia_mask bits of slots past kMaxReactions stay clear)
  switch ((a.nb + 3) / 4) {
    case 1: by_k(std::integral_constant<int, 1>{}); break;
    case 2: by_k(std::integral_constant<int, 2>{}); break;
    case 3: by_k(std::integral_constant<int, 3>{}); break;
    case 4: by_k(std::integral_constant<int, 4>{}); break;
    case 5: by_k(std::integral_constant<int, 5>{}); break;
    case 6: by_k(std::integral_constant<int, 6>{}); break;
    case 7: by_k(std::integral_constant<int, 7>{}); break;
    default: by_k(std::integral_constant<int, 8>{}); break;
  }
  NC_CHECK_HIP(hipGetLastError());
}

}  // namespace

// FP32 storage only: with FP64 stage-input copies the epilogue kernel measured 3% slower
// over the C3 solve than the separate stage passes (8% faster in FP32 storage)
bool sens_rk_fusable(const nc_sensmatrix_s* SA) { return SA->P + 1 <= 16 && SA->A->precision == NC_FP32; }

void sens_matvec_launch(nc_sensmatrix_s* SA, double t, const void* v, int64_t vstride, double* out, int64_t ostride,
                        hipStream_t s, const RkEpilogue* epi) {
  nc_matrix_s* A = SA->A;
  matrix_prepare(A, t);
  if (SA->has_joint && t != SA->t_cache) {
    SA->t_cache = t;
    if (A->precision == NC_FP32)
      fill_pairs<float>(SA, SA->states.p, SA->state_conn.p, t, true, nullptr);
    else
      fill_pairs<double>(SA, SA->states.p, SA->state_conn.p, t, true, nullptr);
  }
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
    launch_sens<float>(SA, t, v, vstride, out, ostride, s, e0, e1, epi);
  else
    launch_sens<double>(SA, t, v, vstride, out, ostride, s, e0, e1, epi);
  if (timed) A->ev_used++;
  A->launches++;
}

nc_sensmatrix_s* sensmatrix_create(nc_space_s* sp, nc_model_s* m, int precision) {
  NC_REQUIRE(m->has_grad, NC_ERR_ARG, "the model has no propensity gradients (CmeModelWithSensitivity)");
  NC_REQUIRE(m->np + 1 <= kMaxSensBlocks, NC_ERR_UNSUPPORTED, "at most 31 parameters");
  auto SA = std::make_unique<nc_sensmatrix_s>();
  SA->A = matrix_create(sp, m, precision);
  SA->model = m;
  SA->P = m->np;
  nc_matrix_s* A = SA->A;
  // pairs of the sparsity pattern, grouped by parameter (block p + 1)
  std::vector<int> cslot_of_r(m->nr, -1);
  int64_t sink_total = 0;
  SA->pstart.assign(SA->P + 2, 0);
  for (int p = 0; p < m->np; ++p) {
    SA->pstart[p + 1] = (int32_t)SA->pairs.size();
    for (int r = 0; r < m->nr; ++r) {
      if (!m->pattern[(size_t)r * m->np + p]) continue;
      SensPair sp_{};
      sp_.r = r;
      sp_.p = p;
      sp_.kind = m->kinds[r];
      sp_.q = -1;
      if (sp_.kind == NC_PROP_SEPARABLE)
        for (size_t q = 0; q < A->tv_r.size(); ++q)
          if (A->tv_r[q] == r) sp_.q = (int32_t)q;
      sp_.arr = -1;
      sp_.cslot = -1;
      if (m->dxlen(r, p) > 0) {
        sp_.arr = SA->n_arr++;
        if (cslot_of_r[r] < 0) cslot_of_r[r] = SA->n_cslots++;
        sp_.cslot = cslot_of_r[r];
        sp_.sink_off = sink_total;
        sink_total += A->sink_off_h[r + 1] - A->sink_off_h[r];
      }
      SA->pairs.push_back(sp_);
    }
  }
  SA->pstart[SA->P + 1] = (int32_t)SA->pairs.size();
  NC_REQUIRE((int)SA->pairs.size() <= kMaxPairs, NC_ERR_UNSUPPORTED, "too many (reaction, parameter) derivative pairs");
  // pstart is indexed by block: block b (b >= 1) holds parameter b - 1
  const size_t vs = A->vsize;
  const int64_t ld = std::max<int64_t>(A->ld, 1);
  SA->ddiag.alloc((size_t)std::max(SA->n_arr, 1) * ld * vs);
  SA->dpred.alloc((size_t)std::max(SA->n_arr, 1) * ld * vs);
  SA->dcol.alloc((size_t)std::max(SA->n_cslots, 1) * ld);
  SA->dsink.alloc((size_t)std::max<int64_t>(sink_total, 1) * vs);
  std::vector<PairMeta> meta(std::max<size_t>(SA->pairs.size(), 1));
  for (size_t k = 0; k < SA->pairs.size(); ++k) {
    const SensPair& p = SA->pairs[k];
    meta[k] = PairMeta{p.r, p.arr, p.cslot, p.q, p.sink_off};
    SA->has_joint = SA->has_joint || (p.kind == NC_PROP_JOINT && p.arr >= 0);
  }
  hipStream_t s = A->stream.s;
  SA->meta_dev.alloc(meta.size() * sizeof(PairMeta));
  NC_CHECK_HIP(hipMemcpyAsync(SA->meta_dev.p, meta.data(), meta.size() * sizeof(PairMeta), hipMemcpyHostToDevice, s));
  // the row part's derivative terms, grouped by the wave that owns the pair's block: the
  // P + 1 blocks go to the four waves in contiguous chunks, the first (P + 1) % 4 waves
  // taking one block more (block b = p + 1 is accumulator b - bstart[w] of its wave w)
  {
    constexpr int W = kBlock / 64;
    const int nb = SA->P + 1, base = nb / W, extra = nb % W;
    SA->bstart[0] = 0;
    for (int w = 0; w < W; ++w) SA->bstart[w + 1] = SA->bstart[w] + base + (w < extra ? 1 : 0);
    std::vector<Term> terms;
    for (int w = 0; w < W; ++w) {
      SA->tstart[w] = (int32_t)terms.size();
      for (size_t k = 0; k < SA->pairs.size(); ++k) {
        const SensPair& p = SA->pairs[k];
        const int bb = p.p + 1;
        if (bb < SA->bstart[w] || bb >= SA->bstart[w + 1]) continue;
        Term T{};
        T.kk = (int8_t)(bb - SA->bstart[w]);
        T.pair = (int16_t)k;
        if (p.arr >= 0) {
          T.kind = 0;
          T.idx = (int16_t)p.arr;
          T.cslot = (int16_t)p.cslot;
          terms.push_back(T);
        }
        if (p.q >= 0) {
          T.kind = 1;
          T.idx = (int16_t)p.q;
          T.cslot = 0;
          terms.push_back(T);
        }
      }
    }
    SA->tstart[W] = (int32_t)terms.size();
    NC_REQUIRE((int)terms.size() <= kMaxTerms, NC_ERR_UNSUPPORTED, "too many derivative terms");
    if (terms.empty()) terms.push_back(Term{0, 0, -1, 0, 0, 0, 0});
    SA->terms_dev.alloc(terms.size() * sizeof(Term));
    NC_CHECK_HIP(hipMemcpyAsync(SA->terms_dev.p, terms.data(), terms.size() * sizeof(Term), hipMemcpyHostToDevice, s));
  }
  // tile fold: when the generator reaches rows farther apart than an XCD's contiguous share
  // of the tiles (a slow species with few levels, such as the gene state of a one-hot
  // model, placed last in lattice order), interleave the segments one reach apart
  {
    const int64_t tiles = (A->n + 63) / 64, reach_tiles = (A->max_reach + 63) / 64;
    SA->fold = 1;
    if (reach_tiles > 0 && reach_tiles >= tiles / 8) {
      const int64_t f = tiles / reach_tiles;
      if (f >= 2 && f <= 64) SA->fold = (int32_t)f;
    }
  }
  if (SA->has_joint && sp->n) {
    SA->states.alloc((size_t)sp->n * sp->ns);
    SA->state_conn.alloc((size_t)sp->n * sp->nr);
    NC_CHECK_HIP(hipMemcpyAsync(SA->states.p, sp->states.p, sizeof(int32_t) * sp->n * sp->ns, hipMemcpyDeviceToDevice, s));
    NC_CHECK_HIP(hipMemcpyAsync(SA->state_conn.p, sp->state_conn.p, sizeof(int32_t) * sp->n * sp->nr,
                                hipMemcpyDeviceToDevice, s));
  }
  DevBuf<unsigned long long> dmax(1);
  NC_CHECK_HIP(hipMemsetAsync(dmax.p, 0, sizeof(unsigned long long), s));
  NC_CHECK_HIP(hipStreamSynchronize(sp->stream.s));
  if (precision == NC_FP32)
    fill_pairs<float>(SA.get(), sp->states.p, sp->state_conn.p, 0.0, false, dmax.p);
  else
    fill_pairs<double>(SA.get(), sp->states.p, sp->state_conn.p, 0.0, false, dmax.p);
  unsigned long long bits = 0;
  NC_CHECK_HIP(hipMemcpyAsync(&bits, dmax.p, sizeof(bits), hipMemcpyDeviceToHost, s));
  NC_CHECK_HIP(hipStreamSynchronize(s));
  double dm;
  std::memcpy(&dm, &bits, sizeof(dm));
  SA->t_cache = SA->has_joint ? 0.0 : -INFINITY;
  SA->norm1 = A->norm1 + 2.0 * dm * (double)std::max<size_t>(SA->pairs.size(), 1);
  return SA.release();
}

}  // namespace nc

using namespace nc;

namespace {
template <typename VT>
__global__ void k_d2vt(const double* in, VT* out, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = (VT)in[i];
}
}  // namespace

extern "C" {

int nc_sensmatrix_create(nc_space sp, nc_model m, int32_t precision, nc_sensmatrix* out) {
  return guard([&] {
    NC_REQUIRE(sp && m && out, NC_ERR_ARG, "null argument");
    *out = sensmatrix_create(sp, m, precision);
  });
}

int nc_sensmatrix_destroy(nc_sensmatrix SA) {
  return guard([&] { delete SA; });
}

int nc_sensmatrix_info(nc_sensmatrix SA, int64_t* n_rows, int32_t* n_params, int32_t* n_pairs, double* bytes) {
  return guard([&] {
    NC_REQUIRE(SA, NC_ERR_ARG, "null argument");
    const nc_matrix_s* A = SA->A;
    if (n_rows) *n_rows = A->n_rows;
    if (n_params) *n_params = SA->P;
    if (n_pairs) *n_pairs = (int32_t)SA->pairs.size();
    if (bytes) {
      // matrix once, the P+1 input blocks (VT) once, the P+1 FP64 output blocks, pair arrays
      const double vs = (double)A->vsize, nb = SA->P + 1.0;
      *bytes = matrix_bytes_per_matvec(A) - A->n_rows * (vs + 8.0) + nb * A->n_rows * (vs + 8.0) +
               SA->n_arr * (double)A->n * 2.0 * vs + SA->n_cslots * (double)A->n * 4.0;
    }
  });
}

int nc_sensmatrix_matvec(nc_sensmatrix SA, double t, const double* v, double* out) {
  return guard([&] {
    NC_REQUIRE(SA && v && out, NC_ERR_ARG, "null argument");
    nc_matrix_s* A = SA->A;
    hipStream_t s = A->stream.s;
    const int64_t N = A->n_rows, tot = N * (SA->P + 1);
    DevBuf<double> dv(std::max<int64_t>(tot, 1)), dout(std::max<int64_t>(tot, 1));
    DevBuf<unsigned char> vt(std::max<int64_t>(tot, 1) * A->vsize);
    NC_CHECK_HIP(hipMemcpyAsync(dv.p, v, sizeof(double) * tot, hipMemcpyHostToDevice, s));
    if (A->precision == NC_FP32)
      k_d2vt<float><<<grid_for(tot, 256), 256, 0, s>>>(dv.p, reinterpret_cast<float*>(vt.p), tot);
    else
      k_d2vt<double><<<grid_for(tot, 256), 256, 0, s>>>(dv.p, reinterpret_cast<double*>(vt.p), tot);
    NC_CHECK_HIP(hipGetLastError());
    sens_matvec_launch(SA, t, vt.p, N, dout.p, N, s);
    NC_CHECK_HIP(hipMemcpyAsync(out, dout.p, sizeof(double) * tot, hipMemcpyDeviceToHost, s));
    NC_CHECK_HIP(hipStreamSynchronize(s));
    matvec_harvest_timing(A);
  });
}

int nc_sensmatrix_bench_matvec(nc_sensmatrix SA, double t, int32_t reps, double* mean_ms) {
  return guard([&] {
    NC_REQUIRE(SA && mean_ms && reps > 0, NC_ERR_ARG, "bad arguments");
    nc_matrix_s* A = SA->A;
    hipStream_t s = A->stream.s;
    const int64_t N = A->n_rows, tot = N * (SA->P + 1);
    DevBuf<unsigned char> vt(std::max<int64_t>(tot, 1) * A->vsize);
    DevBuf<double> o(std::max<int64_t>(tot, 1));
    NC_CHECK_HIP(hipMemsetAsync(vt.p, 0, tot * A->vsize, s));
    const bool was = A->timing;
    A->timing = false;
    sens_matvec_launch(SA, t, vt.p, N, o.p, N, s);
    hipEvent_t e0, e1;
    NC_CHECK_HIP(hipEventCreate(&e0));
    NC_CHECK_HIP(hipEventCreate(&e1));
    NC_CHECK_HIP(hipEventRecord(e0, s));
    for (int k = 0; k < reps; ++k) sens_matvec_launch(SA, t, vt.p, N, o.p, N, s);
    NC_CHECK_HIP(hipEventRecord(e1, s));
    NC_CHECK_HIP(hipEventSynchronize(e1));
    A->timing = was;
    float ms = 0.f;
    NC_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *mean_ms = ms / reps;
  });
}

}  // extern "C"
