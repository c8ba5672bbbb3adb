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

#include "nc_sens.h"

nc_sensmatrix_s::~nc_sensmatrix_s() { delete A; }

namespace nc {
namespace {

constexpr int kBlock = 256;

struct PairMeta {
  int32_t r, arr, cslot, q;
  int64_t sink_off;
};

template <typename VT>
struct SensArgs {
  int64_t n, ld, vstride, ostride;
  int32_t K, ntv, nr, nb;
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
  int32_t pstart[kMaxSensBlocks + 1];  // pairs of block b (b >= 1): pstart[b] .. pstart[b+1)
  double tv_scale[kMaxReactions];
  double sink_scale[kMaxReactions];
  double dscale[kMaxPairs];   // multiplies the pair's state-factor derivative term
  double sfscale[kMaxPairs];  // multiplies the time-factor derivative term (separable)
};

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

template <typename VT, int MB>
__global__ __launch_bounds__(kBlock) void k_sens_matvec(SensArgs<VT> a, const VT* __restrict__ v,
                                                        double* __restrict__ out) {
  const int64_t row_blocks = (a.n + kBlock - 1) / kBlock;
  const int nb = a.nb;
  if ((int64_t)blockIdx.x < row_blocks) {
    // XCD-contiguous row blocks, as in the plain mat-vec
    const int64_t b = blockIdx.x, q8 = row_blocks >> 3, r8 = row_blocks & 7, x = b & 7;
    const int64_t i = (x * q8 + (x < r8 ? x : r8) + (b >> 3)) * kBlock + threadIdx.x;
    if (i >= a.n) return;
    double acc[MB], vi[MB];
    const double dg = a.diag[i];
#pragma unroll
    for (int k = 0; k < MB; ++k) {
      vi[k] = k < nb ? (double)v[k * a.vstride + i] : 0.0;
      acc[k] = dg * vi[k];
    }
    for (int k = 0; k < a.K; ++k) {
      const int32_t c = a.col[k * a.ld + i];
      const double w = (double)a.val[k * a.ld + i];
#pragma unroll
      for (int bb = 0; bb < MB; ++bb)
        if (bb < nb) acc[bb] += w * (double)v[bb * a.vstride + c];
    }
    for (int q = 0; q < a.ntv; ++q) {
      const double s = a.tv_scale[q];
      const double d = (double)a.tvd[q * a.ld + i], w = (double)a.tvv[q * a.ld + i];
      const int32_t c = a.tvc[q * a.ld + i];
#pragma unroll
      for (int bb = 0; bb < MB; ++bb)
        if (bb < nb) acc[bb] += s * (d * vi[bb] + w * (double)v[bb * a.vstride + c]);
    }
    // dA_p v_0 (block p + 1)
#pragma unroll
    for (int bb = 1; bb < MB; ++bb) {
      if (bb >= nb) break;
      double add = 0.0;
      for (int e = a.pstart[bb]; e < a.pstart[bb + 1]; ++e) {
        const PairMeta m = a.meta[e];
        if (m.arr >= 0) {
          const int32_t c = a.dcol[m.cslot * a.ld + i];
          add += a.dscale[e] * ((double)a.ddiag[m.arr * a.ld + i] * vi[0] + (double)a.dpred[m.arr * a.ld + i] * (double)v[c]);
        }
        if (m.q >= 0 && a.sfscale[e] != 0.0) {
          const int q = m.q;
          const int32_t c = a.tvc[q * a.ld + i];
          add += a.sfscale[e] * ((double)a.tvd[q * a.ld + i] * vi[0] + (double)a.tvv[q * a.ld + i] * (double)v[c]);
        }
      }
      acc[bb] += add;
    }
#pragma unroll
    for (int bb = 0; bb < MB; ++bb)
      if (bb < nb) out[bb * a.ostride + i] = acc[bb];
  } else {
    // sink row r of block bb: one workgroup, deterministic tree
    const int sb = (int)((int64_t)blockIdx.x - row_blocks);
    const int bb = sb / a.nr, r = sb % a.nr;
    const int64_t e0 = a.sink_off[r], e1 = a.sink_off[r + 1];
    const VT* vb = v + bb * a.vstride;
    double acc = 0.0;
    for (int64_t e = e0 + threadIdx.x; e < e1; e += kBlock) acc += (double)a.sink_val[e] * (double)vb[a.sink_src[e]];
    acc *= a.sink_scale[r];
    if (bb >= 1) {
      for (int p = a.pstart[bb]; p < a.pstart[bb + 1]; ++p) {
        const PairMeta m = a.meta[p];
        if (m.r != r) continue;
        double s1 = 0.0, s2 = 0.0;
        for (int64_t e = e0 + threadIdx.x; e < e1; e += kBlock) {
          const double v0 = (double)v[a.sink_src[e]];
          if (m.arr >= 0) s1 += (double)a.dsink[m.sink_off + (e - e0)] * v0;
          if (m.q >= 0) s2 += (double)a.sink_val[e] * v0;
        }
        acc += a.dscale[p] * s1 + a.sfscale[p] * s2;
      }
    }
    __shared__ double part[kBlock / 64];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) s += part[w];
      out[bb * a.ostride + a.n + r] = s;
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
                 hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
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
  matrix_scales(A, t, a.sink_scale, a.tv_scale);
  for (int b = 0; b <= SA->P + 1; ++b) a.pstart[b] = SA->pstart[b];
  for (size_t k = 0; k < SA->pairs.size(); ++k) {
    const SensPair& p = SA->pairs[k];
    a.dscale[k] = p.kind == NC_PROP_SEPARABLE ? m->tfactor(p.r, t) : 1.0;
    a.sfscale[k] = (p.q >= 0 && m->dtlen(p.r, p.p) > 0)
                       ? prog_eval(m->dtprog(p.r, p.p), m->dtlen(p.r, p.p), t, nullptr, m->params.data())
                       : 0.0;
  }
  const int64_t row_blocks = (A->n + kBlock - 1) / kBlock;
  const unsigned grid = (unsigned)(row_blocks + (int64_t)a.nb * A->nr);
  const VT* vv = reinterpret_cast<const VT*>(v);
  auto go = [&](auto kern) {
    if (e0)
      hipExtLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, s, e0, e1, 0, a, vv, out);
    else
      kern<<<grid, kBlock, 0, s>>>(a, vv, out);
  };
  if (a.nb <= 2)
    go(k_sens_matvec<VT, 2>);
  else if (a.nb <= 4)
    go(k_sens_matvec<VT, 4>);
  else if (a.nb <= 8)
    go(k_sens_matvec<VT, 8>);
  else if (a.nb <= 16)
    go(k_sens_matvec<VT, 16>);
  else
    go(k_sens_matvec<VT, 32>);
  NC_CHECK_HIP(hipGetLastError());
}

}  // namespace

void sens_matvec_launch(nc_sensmatrix_s* SA, double t, const void* v, int64_t vstride, double* out, int64_t ostride,
                        hipStream_t s) {
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
    launch_sens<float>(SA, t, v, vstride, out, ostride, s, e0, e1);
  else
    launch_sens<double>(SA, t, v, vstride, out, ostride, s, e0, e1);
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
