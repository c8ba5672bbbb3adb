// Learned space adapter: boundary-state features on the device, the network's keep/add
// selection, and the training records gathered from plain SelectiveRStep solves. No
// reference counterpart: NumCME.jl's adapters are rule-based (rstepadapters.jl:55-110).
//
// Feature row of a boundary state x (p = its probability, a_r = propensities at time t):
//   f0..f2  x_s / (1 + max_s x_s) for species 0..2 (max over the current space; 0 if s >= ns)
//   f3      log10(max(p, 1e-30)) / 30
//   f4      log10(1 + sum_r a_r) / 4
//   f5      log10(max(share, 1e-30)) / 30,  share = p sum_{r: x + s_r outside} a_r / sum_r dsinks_r
//   f6      t / tend
//   f7      (number of reactions leading outside the space at x) / nr
// Everything is computed in FP64 and stored in FP32.
#include <algorithm>
#include <cmath>
#include <memory>

#include "nc_learned.h"

namespace nc {
namespace {

constexpr int kBlock = 256;
constexpr int kF = NC_MLP_FEATURES;

struct PropDesc {
  const nc_instr* code;
  const int64_t* offs;
  const double* theta;
  int nr;
  int32_t kind[kMaxReactions];
  double scale[kMaxReactions];  // separable time factor at t (1 otherwise)
};

__device__ inline double prop_at(const PropDesc& d, int r, double t, const int32_t* x) {
  const int64_t o = d.offs[2 * r];
  const double v = prog_eval(d.code + o, d.offs[2 * r + 1] - o, t, x, d.theta);
  return d.kind[r] == NC_PROP_JOINT ? v : d.scale[r] * v;
}

__global__ void k_flag_boundary(const uint32_t* sink_mask, int64_t n, uint32_t rmask, uint32_t* flags) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) flags[i] = (sink_mask[i] & rmask) ? 1u : 0u;
}

__global__ void k_scatter_ids(const uint32_t* flags, const uint32_t* ranks, int64_t n, int32_t* E) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n && flags[i]) E[ranks[i]] = (int32_t)i;
}

// per-species maximum of the first three species: grid-stride maxima, wavefront shuffles,
// the block's waves through LDS, and one atomic per block and species (states are
// non-negative, xmax starts at 0; the grid is kept small so few atomics meet one address)
__global__ __launch_bounds__(kBlock) void k_species_max(const int32_t* states, int64_t n, int ns, int32_t* xmax) {
  __shared__ int32_t wm[kBlock / 64][3];
  int32_t m[3] = {0, 0, 0};
  const int nsm = ns < 3 ? ns : 3;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    for (int s = 0; s < nsm; ++s) m[s] = max(m[s], states[i * ns + s]);
  for (int s = 0; s < 3; ++s) {
    for (int off = 32; off > 0; off >>= 1) m[s] = max(m[s], __shfl_xor(m[s], off, 64));
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6][s] = m[s];
  }
  __syncthreads();
  if ((int)threadIdx.x < nsm) {
    int32_t v = 0;
    for (int w = 0; w < kBlock / 64; ++w) v = max(v, wm[w][threadIdx.x]);
    atomicMax(xmax + threadIdx.x, v);
  }
}

__global__ void k_features(int64_t nE, const int32_t* E, const int32_t* states, const uint32_t* sink_mask, int ns,
                           const double* p, PropDesc d, double t, double t_frac, const int32_t* xmax,
                           double inv_flux, float* feat) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= nE) return;
  const int64_t i = E[j];
  int32_t x[kMaxSpecies];
  for (int s = 0; s < ns; ++s) x[s] = states[i * ns + s];
  const uint32_t sm = sink_mask[i];
  double atot = 0.0, asink = 0.0;
  for (int r = 0; r < d.nr; ++r) {
    const double a = prop_at(d, r, t, x);
    atot += a;
    if ((sm >> r) & 1u) asink += a;
  }
  const double pi = p[i];
  float* f = feat + j * kF;
  for (int s = 0; s < 3; ++s) f[s] = s < ns ? (float)((double)x[s] / (1.0 + (double)xmax[s])) : 0.0f;
  f[3] = (float)(log10(fmax(pi, 1e-30)) / 30.0);
  f[4] = (float)(log10(1.0 + atot) / 4.0);
  f[5] = (float)(log10(fmax(pi * asink * inv_flux, 1e-30)) / 30.0);
  f[6] = (float)t_frac;
  f[7] = (float)((double)__popc(sm) / (double)d.nr);
}

__global__ void k_select(int64_t nE, const int32_t* E, const float* logits, float thr, uint32_t* select,
                         uint32_t* counter) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= nE) return;
  if (logits[j] >= thr) {
    select[E[j]] = 1u;
    atomicAdd(counter, 1u);
  }
}

__global__ void k_fill0(uint32_t* p, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// queue rows j = k * stride (k < cnt) of the boundary list as pending records
__global__ void k_records_add(int64_t cnt, int64_t stride, const int32_t* E, const float* feat, const int32_t* states,
                              const uint32_t* sink_mask, uint32_t rmask, int ns, float* rfeat, int32_t* pstates,
                              uint32_t* pmask) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= cnt) return;
  const int64_t j = k * stride;
  const int64_t i = E[j];
  for (int f = 0; f < kF; ++f) rfeat[k * kF + f] = feat[j * kF + f];
  for (int s = 0; s < ns; ++s) pstates[k * ns + s] = states[i * ns + s];
  pmask[k] = sink_mask[i] & rmask;
}

// label = probability of the successors x + s_r (r in the record's mask) now >= label_mass
__global__ void k_records_label(int64_t cnt, const int32_t* pstates, const uint32_t* pmask, const int32_t* stoich,
                                int ns, int nr, const uint32_t* table, uint64_t tmask, const int32_t* states,
                                const double* p, double label_mass, float* labels) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= cnt) return;
  int32_t y[kMaxSpecies];
  double mass = 0.0;
  const uint32_t m = pmask[k];
  for (int r = 0; r < nr; ++r) {
    if (!((m >> r) & 1u)) continue;
    bool nonneg = true;
    for (int s = 0; s < ns; ++s) {
      y[s] = pstates[k * ns + s] + stoich[s + r * ns];
      nonneg = nonneg && y[s] >= 0;
    }
    if (!nonneg) continue;
    const int32_t idx = ht_find(table, tmask, states, ns, y);
    if (idx >= 0) mass += p[idx];
  }
  labels[k] = mass >= label_mass ? 1.0f : 0.0f;
}

}  // namespace

int64_t learned_features(nc_space_s* sp, const nc_model_s* m, const double* p, const double* dsinks, uint32_t rmask,
                         double t, double tend, LearnWork& w, hipStream_t s) {
  const int64_t n = sp->n;
  if (n == 0) return 0;
  NC_REQUIRE(m->ns == sp->ns && m->nr == sp->nr, NC_ERR_ARG, "model and space differ in species or reactions");
  w.flags.reserve(n);
  w.ranks.reserve(n);
  w.E.reserve(n);
  k_flag_boundary<<<grid_for(n, kBlock), kBlock, 0, s>>>(sp->sink_mask.p, n, rmask, w.flags.p);
  uint32_t nE = 0;
  exclusive_scan_u32(w.flags.p, w.ranks.p, n, w.scan_tmp, s, &nE);
  if (nE == 0) return 0;
  k_scatter_ids<<<grid_for(n, kBlock), kBlock, 0, s>>>(w.flags.p, w.ranks.p, n, w.E.p);
  w.xmax.reserve(3);
  NC_CHECK_HIP(hipMemsetAsync(w.xmax.p, 0, sizeof(int32_t) * 3, s));
  k_species_max<<<std::min<unsigned>(grid_for(n, kBlock), 256), kBlock, 0, s>>>(sp->states.p, n, sp->ns, w.xmax.p);
  PropDesc d{};
  d.code = m->d_code.p;
  d.offs = m->d_offs.p;
  d.theta = m->d_params.p;
  d.nr = m->nr;
  for (int r = 0; r < m->nr; ++r) {
    d.kind[r] = m->kinds[r];
    d.scale[r] = m->kinds[r] == NC_PROP_SEPARABLE ? m->tfactor(r, t) : 1.0;
  }
  double flux = 0.0;
  for (int r = 0; r < m->nr; ++r) flux += dsinks[r];
  const double inv_flux = flux > 0.0 ? 1.0 / flux : 0.0;
  w.feat.reserve((size_t)nE * kF);
  k_features<<<grid_for(nE, kBlock), kBlock, 0, s>>>(nE, w.E.p, sp->states.p, sp->sink_mask.p, sp->ns, p, d, t,
                                                     t / tend, w.xmax.p, inv_flux, w.feat.p);
  NC_CHECK_HIP(hipGetLastError());
  return nE;
}

int64_t learned_select(nc_space_s* sp, nc_mlp_s* M, int64_t nE, double threshold, LearnWork& w, hipStream_t s) {
  const int64_t n = sp->n;
  w.select.reserve(std::max<int64_t>(n, 1));
  w.counter.reserve(1);
  k_fill0<<<grid_for(n, kBlock), kBlock, 0, s>>>(w.select.p, n);
  NC_CHECK_HIP(hipMemsetAsync(w.counter.p, 0, sizeof(uint32_t), s));
  if (nE == 0 || !(threshold < 1.0)) {
    NC_CHECK_HIP(hipStreamSynchronize(s));
    return 0;
  }
  // sigmoid(z) >= threshold  <=>  z >= log(threshold / (1 - threshold))
  const float thr = threshold <= 0.0 ? -INFINITY : (float)std::log(threshold / (1.0 - threshold));
  w.logits.reserve(nE);
  mlp_forward_device(M, nE, w.feat.p, w.logits.p, s);
  k_select<<<grid_for(nE, kBlock), kBlock, 0, s>>>(nE, w.E.p, w.logits.p, thr, w.select.p, w.counter.p);
  NC_CHECK_HIP(hipGetLastError());
  uint32_t* hk = static_cast<uint32_t*>(host_stage(5));
  NC_CHECK_HIP(hipMemcpyAsync(hk, w.counter.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  NC_CHECK_HIP(hipStreamSynchronize(s));
  return *hk;
}

void records_label(nc_records_s* R, nc_space_s* sp, const double* p, hipStream_t s) {
  if (!R || R->n_pending == 0) return;
  if (sp->n > 0 && sp->ns == R->ns) {
    k_records_label<<<grid_for(R->n_pending, kBlock), kBlock, 0, s>>>(
        R->n_pending, R->pend_states.p, R->pend_mask.p, R->pend_stoich.p, R->ns, R->pend_nr, sp->table.p,
        sp->table_cap - 1, sp->states.p, p, R->label_mass, R->labels.p + R->n_labelled);
  } else {
    k_fill0<<<grid_for(R->n_pending, kBlock), kBlock, 0, s>>>(reinterpret_cast<uint32_t*>(R->labels.p + R->n_labelled),
                                                             R->n_pending);  // 0.0f has all bits clear
  }
  NC_CHECK_HIP(hipGetLastError());
  NC_CHECK_HIP(hipStreamSynchronize(s));
  R->n_labelled += R->n_pending;
  R->n_pending = 0;
}

void records_add(nc_records_s* R, nc_space_s* sp, int64_t nE, uint32_t rmask, LearnWork& w, hipStream_t s) {
  if (!R || nE == 0) return;
  NC_REQUIRE(R->n_pending == 0, NC_ERR_INTERNAL, "training records: previous records not labelled");
  const int64_t room = R->capacity - R->n_labelled;
  if (room <= 0) return;
  const int64_t stride = (nE + R->per_adaptation - 1) / R->per_adaptation;
  const int64_t cnt = std::min<int64_t>((nE + stride - 1) / stride, room);
  R->ns = sp->ns;
  R->pend_nr = sp->nr;
  R->pend_states.reserve((size_t)cnt * sp->ns);
  R->pend_mask.reserve(cnt);
  R->pend_stoich.reserve((size_t)sp->ns * sp->nr);
  NC_CHECK_HIP(hipMemcpyAsync(R->pend_stoich.p, sp->d_stoich.p, sizeof(int32_t) * sp->ns * sp->nr,
                              hipMemcpyDeviceToDevice, s));
  k_records_add<<<grid_for(cnt, kBlock), kBlock, 0, s>>>(cnt, stride, w.E.p, w.feat.p, sp->states.p, sp->sink_mask.p,
                                                         rmask, sp->ns, R->features.p + R->n_labelled * kF,
                                                         R->pend_states.p, R->pend_mask.p);
  NC_CHECK_HIP(hipGetLastError());
  R->n_pending = cnt;
}

}  // namespace nc

using namespace nc;

extern "C" {

int nc_adapter_features(nc_space sp, nc_model model, int64_t n, const double* p, const double* dsinks, double t,
                        double tend, int64_t* n_boundary, int64_t* ids, float* features) {
  return guard([&] {
    NC_REQUIRE(sp && model && n_boundary && dsinks, NC_ERR_ARG, "null argument");
    NC_REQUIRE(n == sp->n, NC_ERR_ARG, "p must hold one probability per state");
    NC_REQUIRE(n == 0 || (p && ids && features), NC_ERR_ARG, "null argument");
    NC_REQUIRE(tend > 0.0, NC_ERR_ARG, "tend must be positive");
    *n_boundary = 0;
    if (n == 0) return;
    uint32_t rmask = 0;
    for (int r = 0; r < sp->nr; ++r)
      if (dsinks[r] > 0.0) rmask |= 1u << r;
    if (rmask == 0) rmask = sp->nr >= 32 ? 0xFFFFFFFFu : ((1u << sp->nr) - 1u);
    hipStream_t s = sp->stream.s;
    DevBuf<double> dp(n);
    NC_CHECK_HIP(hipMemcpyAsync(dp.p, p, sizeof(double) * n, hipMemcpyHostToDevice, s));
    LearnWork w;
    const int64_t nE = learned_features(sp, model, dp.p, dsinks, rmask, t, tend, w, s);
    if (nE) {
      std::vector<int32_t> E(nE);
      NC_CHECK_HIP(hipMemcpyAsync(E.data(), w.E.p, sizeof(int32_t) * nE, hipMemcpyDeviceToHost, s));
      NC_CHECK_HIP(hipMemcpyAsync(features, w.feat.p, sizeof(float) * nE * kF, hipMemcpyDeviceToHost, s));
      NC_CHECK_HIP(hipStreamSynchronize(s));
      for (int64_t j = 0; j < nE; ++j) ids[j] = E[j];
    }
    NC_CHECK_HIP(hipStreamSynchronize(s));
    *n_boundary = nE;
  });
}

int nc_records_create(int64_t capacity, int64_t per_adaptation, double label_mass, nc_records* out) {
  return guard([&] {
    NC_REQUIRE(out && capacity >= 1 && per_adaptation >= 1, NC_ERR_ARG, "bad arguments");
    NC_REQUIRE(label_mass > 0.0 && std::isfinite(label_mass), NC_ERR_ARG, "label_mass must be positive");
    auto R = std::make_unique<nc_records_s>();
    R->capacity = capacity;
    R->per_adaptation = per_adaptation;
    R->label_mass = label_mass;
    R->features.alloc((size_t)capacity * kF);
    R->labels.alloc((size_t)capacity);
    *out = R.release();
  });
}

int nc_records_destroy(nc_records R) {
  return guard([&] { delete R; });
}

int nc_records_count(nc_records R, int64_t* n) {
  return guard([&] {
    NC_REQUIRE(R && n, NC_ERR_ARG, "null argument");
    *n = R->n_labelled;
  });
}

int nc_records_get(nc_records R, float* features, float* labels) {
  return guard([&] {
    NC_REQUIRE(R, NC_ERR_ARG, "null argument");
    if (R->n_labelled == 0) return;
    if (features)
      NC_CHECK_HIP(hipMemcpy(features, R->features.p, sizeof(float) * R->n_labelled * kF, hipMemcpyDeviceToHost));
    if (labels) NC_CHECK_HIP(hipMemcpy(labels, R->labels.p, sizeof(float) * R->n_labelled, hipMemcpyDeviceToHost));
  });
}

int nc_records_clear(nc_records R) {
  return guard([&] {
    NC_REQUIRE(R, NC_ERR_ARG, "null argument");
    R->n_labelled = 0;
    R->n_pending = 0;
  });
}

}  // extern "C"
