// Adaptive FSP: the RStep / SelectiveRStep space adapters and the adapt-and-restart
// driver of the adaptive solve.
//
// Drop selection (rstepadapters.jl:41-43, 94-96) runs on the device:
//   keys   = order-preserving uint64 image of each probability (IEEE total order, so
//            -0.0 sorts before +0.0 as Julia's isless does)
//   perm   = stable radix sort of the keys (Julia's sortperm is stable)
//   cs     = inclusive prefix sum of p[perm] in FP64 with a fixed association order:
//            1024-element tiles scanned Hillis-Steele style (step d = 1, 2, ..., 512:
//            x[i] <- x[i] + x[i-d]), tile totals scanned recursively the same way and
//            the prefix of the preceding tiles added to each element
//   total  = cs[n-1] (the reference's sum(p); the two differ by rounding only)
//   drop   = #{k : total - cs[k] >= thr} (">" for the selective adapter),
//            thr = 1 - t*fsptol/tend; the first `drop` entries of perm are removed.
// The fixed association order makes the set bit-reproducible by a vectorised CPU
// restatement given the same p.
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <memory>
#include <mutex>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "nc_fspvec.h"
#include "nc_learned.h"
#include "nc_solver.h"

namespace nc {
namespace {

constexpr int kBlock = 256;
constexpr int kTile = 1024;

__global__ void k_drop_keys(const double* __restrict__ p, int64_t n, uint64_t* __restrict__ keys,
                            uint32_t* __restrict__ idx, unsigned* bad) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = p[i];
  if (!(x == x)) atomicOr(bad, 1u);  // NaN has no place in the order
  const uint64_t u = (uint64_t)__double_as_longlong(x);
  keys[i] = (u >> 63) ? ~u : (u | 0x8000000000000000ull);
  idx[i] = (uint32_t)i;
}

__global__ void k_gather_sorted(const double* __restrict__ p, const uint32_t* __restrict__ perm, int64_t n,
                                double* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = p[perm[i]];
}

// Inclusive Hillis-Steele scan of one 1024-element tile per block.
__global__ __launch_bounds__(kTile) void k_tile_scan(const double* __restrict__ in, double* __restrict__ out,
                                                     double* __restrict__ tile_total, int64_t n) {
  __shared__ double sh[kTile];
  const int tid = threadIdx.x;
  const int64_t i = (int64_t)blockIdx.x * kTile + tid;
  double v = i < n ? in[i] : 0.0;
  sh[tid] = v;
  __syncthreads();
  for (int d = 1; d < kTile; d <<= 1) {
    const double add = tid >= d ? sh[tid - d] : 0.0;
    __syncthreads();
    if (tid >= d) v = v + add;
    sh[tid] = v;
    __syncthreads();
  }
  if (i < n) out[i] = v;
  if (tile_total && tid == kTile - 1) tile_total[blockIdx.x] = v;
}

__global__ void k_add_tile_prefix(double* __restrict__ out, const double* __restrict__ prefix, int64_t n) {
  const int64_t i = (int64_t)kTile + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = out[i] + prefix[i / kTile - 1];
}

__global__ void k_count_drop(const double* __restrict__ cs, int64_t n, double thr, int strict,
                             unsigned long long* count) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const double total = cs[n - 1];
  int hit = 0;
  if (i < n) {
    const double rest = total - cs[i];
    hit = strict ? (rest > thr) : (rest >= thr);
  }
  const unsigned long long b = __ballot(hit);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(count, (unsigned long long)__popcll(b));
}

__global__ void k_u32_to_i64(const uint32_t* __restrict__ in, int64_t n, int64_t* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = (int64_t)in[i];
}

__global__ void k_d2f(const double* __restrict__ in, int64_t n, float* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)in[i];
}

// Inclusive scan with the tile association order described above. Level l holds the
// tile totals of level l-1; tot[l] / pre[l] are their values before / after the scan.
// All level buffers are sized before the first launch, so no pointer moves later.
void scan_fixed(const double* in, double* out, int64_t n, hipStream_t s, std::vector<DevBuf<double>>& tmp) {
  std::vector<int64_t> size{n};
  while (size.back() > kTile) size.push_back((size.back() + kTile - 1) / kTile);
  const size_t levels = size.size();
  if (tmp.size() < 2 * levels) tmp.resize(2 * levels);
  for (size_t l = 1; l < levels; ++l) {
    tmp[2 * l].reserve(size[l]);
    tmp[2 * l + 1].reserve(size[l]);
  }
  auto tot = [&](size_t l) -> double* { return l == 0 ? nullptr : tmp[2 * l].p; };
  auto pre = [&](size_t l) -> double* { return l == 0 ? out : tmp[2 * l + 1].p; };
  // tile scans, finest level first; each writes its tile totals one level up
  for (size_t l = 0; l < levels; ++l) {
    const double* src = l == 0 ? in : tot(l);
    const int64_t tiles = (size[l] + kTile - 1) / kTile;
    k_tile_scan<<<(unsigned)tiles, kTile, 0, s>>>(src, pre(l), l + 1 < levels ? tot(l + 1) : nullptr, size[l]);
    NC_CHECK_HIP(hipGetLastError());
  }
  // prefixes of the preceding tiles, coarsest level first
  for (size_t l = levels - 1; l-- > 0;) {
    k_add_tile_prefix<<<grid_for(size[l] - kTile, kBlock), kBlock, 0, s>>>(pre(l), pre(l + 1), size[l]);
    NC_CHECK_HIP(hipGetLastError());
  }
}

struct DropWork {
  DevBuf<uint64_t> keys, keys_sorted;
  DevBuf<uint32_t> idx, perm, ids32;
  DevBuf<double> sorted, cs;
  std::vector<DevBuf<double>> scan_tmp;
  DevBuf<unsigned char> sort_tmp;
  DevBuf<unsigned long long> counter;
  DevBuf<unsigned> bad;
};

// Sorted 0-based ids of the states the adapter drops; returns their number.
int64_t select_drop(const double* p, int64_t n, double thr, bool strict, hipStream_t s, DropWork& w,
                    DevBuf<int64_t>& ids) {
  if (n <= 0) return 0;
  NC_REQUIRE(n < (int64_t)0x7FFFFFFF, NC_ERR_OVERFLOW, "too many states for the drop selection");
  w.keys.reserve(n);
  w.keys_sorted.reserve(n);
  w.idx.reserve(n);
  w.perm.reserve(n);
  w.sorted.reserve(n);
  w.cs.reserve(n);
  w.counter.reserve(1);
  w.bad.reserve(1);
  NC_CHECK_HIP(hipMemsetAsync(w.counter.p, 0, sizeof(unsigned long long), s));
  NC_CHECK_HIP(hipMemsetAsync(w.bad.p, 0, sizeof(unsigned), s));
  const unsigned g = grid_for(n, kBlock);
  k_drop_keys<<<g, kBlock, 0, s>>>(p, n, w.keys.p, w.idx.p, w.bad.p);
  NC_CHECK_HIP(hipGetLastError());
  size_t bytes = 0;
  NC_CHECK_HIP(rocprim::radix_sort_pairs(nullptr, bytes, w.keys.p, w.keys_sorted.p, w.idx.p, w.perm.p, (size_t)n, 0,
                                         64, s));
  w.sort_tmp.reserve(bytes + 16);
  NC_CHECK_HIP(rocprim::radix_sort_pairs(w.sort_tmp.p, bytes, w.keys.p, w.keys_sorted.p, w.idx.p, w.perm.p, (size_t)n,
                                         0, 64, s));
  k_gather_sorted<<<g, kBlock, 0, s>>>(p, w.perm.p, n, w.sorted.p);
  NC_CHECK_HIP(hipGetLastError());
  scan_fixed(w.sorted.p, w.cs.p, n, s, w.scan_tmp);
  k_count_drop<<<g, kBlock, 0, s>>>(w.cs.p, n, thr, strict ? 1 : 0, w.counter.p);
  NC_CHECK_HIP(hipGetLastError());
  unsigned long long* hm = static_cast<unsigned long long*>(host_stage(4));
  unsigned* hbad = reinterpret_cast<unsigned*>(hm + 1);
  NC_CHECK_HIP(hipMemcpyAsync(hm, w.counter.p, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  NC_CHECK_HIP(hipMemcpyAsync(hbad, w.bad.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
  NC_CHECK_HIP(hipStreamSynchronize(s));
  const unsigned long long m = *hm;
  const unsigned bad = *hbad;
  NC_REQUIRE(!bad, NC_ERR_NUMERIC, "NaN in the probability vector at a space adaptation");
  if (m == 0) return 0;
  // dropids = sort(pids[1:dropcount])
  w.ids32.reserve(m);
  bytes = 0;
  NC_CHECK_HIP(rocprim::radix_sort_keys(nullptr, bytes, w.perm.p, w.ids32.p, (size_t)m, 0, 32, s));
  w.sort_tmp.reserve(bytes + 16);
  NC_CHECK_HIP(rocprim::radix_sort_keys(w.sort_tmp.p, bytes, w.perm.p, w.ids32.p, (size_t)m, 0, 32, s));
  ids.reserve(m);
  k_u32_to_i64<<<grid_for((int64_t)m, kBlock), kBlock, 0, s>>>(w.ids32.p, (int64_t)m, ids.p);
  NC_CHECK_HIP(hipGetLastError());
  return (int64_t)m;
}

void check_adapter(const nc_adapter* ad) {
  NC_REQUIRE(ad, NC_ERR_ARG, "null adapter");
  NC_REQUIRE(ad->kind == NC_ADAPTER_RSTEP || ad->kind == NC_ADAPTER_SELECTIVE_RSTEP ||
                 ad->kind == NC_ADAPTER_LEARNED,
             NC_ERR_ARG, "unknown adapter kind");
  NC_REQUIRE(ad->initial_step_count >= 0 && ad->max_step_count >= 0, NC_ERR_ARG, "step counts must be non-negative");
  if (ad->kind == NC_ADAPTER_LEARNED) {
    NC_REQUIRE(ad->mlp, NC_ERR_ARG, "the learned adapter needs a network");
    NC_REQUIRE(std::isfinite(ad->threshold) && std::isfinite(ad->fallback_dt), NC_ERR_ARG,
               "threshold and fallback_dt must be finite");
  }
}

// SelectiveRStep and the learned adapter built on it expand only along reactions with a
// positive sink derivative and drop with ">".
// Worker threads of nc_solve_adaptive_batch by default: one per hardware queue HIP gives a
// process by default, so that the workers' launches do not queue behind each other.
constexpr int kBatchWorkers = 4;

bool selective_kind(int kind) { return kind != NC_ADAPTER_RSTEP; }

std::vector<int32_t> all_reactions(int nr) {
  std::vector<int32_t> rl(nr);
  for (int r = 0; r < nr; ++r) rl[r] = r;
  return rl;
}

// Reactions whose sink derivative is positive (rstepadapters.jl:104-105); an empty list
// means every reaction to expand! (sparsestatespace.jl:163).
std::vector<int32_t> selective_reactions(const double* dsinks, int nr) {
  std::vector<int32_t> rl;
  for (int r = 0; r < nr; ++r)
    if (dsinks[r] > 0.0) rl.push_back(r);
  if (rl.empty()) rl = all_reactions(nr);
  return rl;
}

}  // namespace
}  // namespace nc

// ---- adaptive solve output ----------------------------------------------------------------
struct nc_output_s {
  int ns = 0, nr = 0;
  int P = 0;  // sensitivity parameters (each snapshot holds P+1 blocks of n + nr)
  struct Space {
    int64_t n = 0;
    std::shared_ptr<nc::DevBuf<int32_t>> states;  // shared with the snapshot vectors
  };
  struct Snap {
    double t = 0.0;
    int32_t space = 0;
    nc::DevBuf<double> u;  // [p; sinks] (and [S_i; dsinks_i] blocks for sensitivities)
  };
  std::vector<std::unique_ptr<Space>> spaces;
  std::vector<std::unique_ptr<Snap>> snaps;
};

using namespace nc;

namespace {

void output_add_space(nc_output_s* o, nc_space_s* sp) {
  auto S = std::make_unique<nc_output_s::Space>();
  S->n = sp->n;
  S->states = share_states(sp->states.p, sp->n, sp->ns, sp->stream.s);
  NC_CHECK_HIP(hipStreamSynchronize(sp->stream.s));
  o->spaces.push_back(std::move(S));
}

// Snapshot of the solver's current state, stored with block stride Nb (unpadded).
void output_add_snap(nc_output_s* o, double t, const nc_solver_s* S) {
  auto P = std::make_unique<nc_output_s::Snap>();
  P->t = t;
  P->space = (int32_t)o->spaces.size() - 1;
  P->u.alloc((size_t)std::max<int64_t>(S->N, 1));
  NC_CHECK_HIP(hipMemcpy2DAsync(P->u.p, sizeof(double) * S->Nb, S->u.p, sizeof(double) * S->Npb,
                                sizeof(double) * S->Nb, S->blocks, hipMemcpyDeviceToDevice, S->s));
  o->snaps.push_back(std::move(P));
}

void accumulate_stats(nc_solve_stats& tot, const nc_solver_s* S) {
  tot.steps += S->st.steps;
  tot.rejected += S->st.rejected;
  tot.matvecs += S->st.matvecs;
  tot.host_syncs += S->st.host_syncs;
  tot.error_estimate += S->st.error_estimate;
  tot.matvecs_timed += S->st.matvecs_timed;
  tot.matvec_ms += S->st.matvec_ms;
}

// Sink rows of du = A(t) u (get_du! of the integrator, rstepadapters.jl:101-104).
std::vector<double> sink_derivative(nc_matrix_s* A, double t, const double* u, hipStream_t s) {
  const int64_t N = A->n_rows;
  DevBuf<double> du(N);
  if (A->precision == NC_FP32) {
    DevBuf<float> v(N);
    k_d2f<<<grid_for(N, kBlock), kBlock, 0, s>>>(u, N, v.p);
    NC_CHECK_HIP(hipGetLastError());
    matvec_launch(A, t, v.p, du.p, false, s, true);
    NC_CHECK_HIP(hipStreamSynchronize(s));
  } else {
    matvec_launch(A, t, u, du.p, false, s, true);
  }
  double* hds = static_cast<double*>(host_stage(4));
  NC_CHECK_HIP(hipMemcpyAsync(hds, du.p + A->n, sizeof(double) * A->nr, hipMemcpyDeviceToHost, s));
  NC_CHECK_HIP(hipStreamSynchronize(s));
  std::vector<double> ds(hds, hds + A->nr);
  matvec_harvest_timing(A);
  return ds;
}

// adapt!: optional drop (the kept probabilities are compacted into p_kept), then expand!.
// p is a device vector of the space's n states. Returns the number dropped; the sorted
// dropped ids are left in `ids` when requested.
// vecs[0] is p; further vectors (sensitivities) are dropped alongside it
// (fsspaceadapterssparse.jl:42-58). kept[b] receives the compacted vecs[b].
// Extras of an adaptation inside the adaptive solve: the learned adapter's network input
// and the training records.
struct AdaptCtx {
  const nc_model_s* model = nullptr;
  nc_records_s* records = nullptr;
  LearnWork* lw = nullptr;
  bool force_plain = false;  // learned kind: adapt as plain SelectiveRStep this time
  bool learned = false;      // out: the network chose the first-level states
};

int64_t adapt_space(nc_space_s* sp, const nc_adapter& ad, const std::vector<const double*>& vecs,
                    const double* dsinks, double t, double tend, double fsptol, DropWork& w, DevBuf<int64_t>& ids,
                    std::vector<DevBuf<double>>& kept, AdaptCtx* ctx = nullptr) {
  hipStream_t s = sp->stream.s;
  const int64_t n = sp->n;
  const double* p = vecs[0];
  int64_t m = 0;
  if (ad.dropstates) {
    const double thr = 1.0 - t * fsptol / tend;
    m = select_drop(p, n, thr, selective_kind(ad.kind), s, w, ids);
  }
  kept.resize(vecs.size());
  for (auto& k : kept) k.reserve((size_t)std::max<int64_t>(n - m, 1));
  if (m) {
    std::vector<double*> outs;
    for (auto& k : kept) outs.push_back(k.p);
    space_delete_device(sp, ids.p, m, (int)vecs.size(), vecs.data(), outs.data());
  } else if (n) {
    for (size_t b = 0; b < vecs.size(); ++b)
      NC_CHECK_HIP(hipMemcpyAsync(kept[b].p, vecs[b], sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  }
  const std::vector<int32_t> rl =
      selective_kind(ad.kind) ? selective_reactions(dsinks, sp->nr) : all_reactions(sp->nr);
  const uint32_t* select = nullptr;
  const bool want_learned = ctx && ad.kind == NC_ADAPTER_LEARNED && !ctx->force_plain;
  const bool want_records = ctx && ctx->records;
  if ((want_learned || want_records) && sp->n > 0) {
    uint32_t rmask = 0;
    for (int32_t r : rl) rmask |= 1u << r;
    const int64_t nE = learned_features(sp, ctx->model, kept[0].p, dsinks, rmask, t, tend, *ctx->lw, s);
    if (want_records) records_add(ctx->records, sp, nE, rmask, *ctx->lw, s);
    if (want_learned && nE > 0 && learned_select(sp, ad.mlp, nE, ad.threshold, *ctx->lw, s) > 0) {
      select = ctx->lw->select.p;
      ctx->learned = true;
    }
  }
  space_expand_list(sp, ad.max_step_count, rl, select);
  NC_CHECK_HIP(hipStreamSynchronize(s));
  return m;
}

// println("t = $(round(tnow, digits=2)). Update state space. New size: $(n).")
// (fspsolve.jl:184-186); Julia prints the rounded Float64 without trailing zeros.
void print_update(double t, int64_t n) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.2f", std::round(t * 100.0) / 100.0);
  std::string v(buf);
  while (v.size() > 2 && v.back() == '0' && v[v.size() - 2] != '.') v.pop_back();
  std::printf("t = %s. Update state space. New size: %lld.\n", v.c_str(), (long long)n);
  std::fflush(stdout);
}

bool is_time_varying(const nc_model_s* m) {
  for (int r = 0; r < m->nr; ++r)
    if (m->kinds[r] != NC_PROP_TIME_INVARIANT) return true;
  return false;
}

// The adaptive solve (fspsolve.jl:105-197), and with P > 0 the adaptive sensitivity solve
// (forwardsenscmesparse.jl:99-215) whose state holds P+1 blocks [p; sinks], [S_i; dsinks_i].
void solve_adaptive_impl(nc_model_s* model, int P, int64_t n0, const int32_t* states0, const double* p0,
                         const double* S0, double ta, double tb, int32_t n_save, const double* saveat,
                         const nc_adaptive_opts& o, nc_output_s* out, nc_solve_stats* st_out) {
  const int ns = model->ns, nr = model->nr, B = P + 1;
  const double tstart = std::min(ta, tb), tend = std::max(ta, tb);
  NC_REQUIRE(std::isfinite(tstart) && std::isfinite(tend), NC_ERR_ARG, "tspan must be finite");
  NC_REQUIRE(o.fsptol > 0.0 && std::isfinite(o.fsptol), NC_ERR_ARG, "fsptol must be positive");
  NC_REQUIRE(n_save >= 0 && (n_save == 0 || saveat), NC_ERR_ARG, "bad saveat");
  for (int k = 1; k < n_save; ++k) NC_REQUIRE(saveat[k] >= saveat[k - 1], NC_ERR_ARG, "saveat must be ascending");
  NC_REQUIRE(n0 >= 1 && states0 && p0, NC_ERR_ARG, "empty initial distribution");
  NC_REQUIRE(P == 0 || S0, NC_ERR_ARG, "null initial sensitivities");
  check_adapter(&o.adapter);
  NC_REQUIRE(P == 0 || o.adapter.kind == NC_ADAPTER_RSTEP, NC_ERR_ARG,
             "sensitivity solves adapt with the R-step rule (ForwardSensRStepAdapter)");
  out->ns = ns;
  out->nr = nr;
  out->P = P;

  // statespace = StateSpaceSparse(stoich, p0.states); init!(...) (fspsolve.jl:122-125)
  std::unique_ptr<nc_space_s> sp(space_create(ns, nr, model->stoich.data(), n0, states0));
  NC_REQUIRE(sp->n == n0, NC_ERR_ARG, "initial distribution states must be unique and non-negative");
  space_expand_list(sp.get(), o.adapter.initial_step_count, all_reactions(nr));
  NC_CHECK_HIP(hipStreamSynchronize(sp->stream.s));

  // unow = [p0; zeros; zeros(sink_count)] per block (forwardsenscmesparse.jl:132-138)
  int64_t Nb = sp->n + nr;
  DevBuf<double> ucur((size_t)(B * Nb));
  NC_CHECK_HIP(hipMemsetAsync(ucur.p, 0, sizeof(double) * B * Nb, sp->stream.s));
  NC_CHECK_HIP(hipMemcpyAsync(ucur.p, p0, sizeof(double) * n0, hipMemcpyHostToDevice, sp->stream.s));
  for (int b = 1; b < B; ++b)
    NC_CHECK_HIP(hipMemcpyAsync(ucur.p + b * Nb, S0 + (size_t)(b - 1) * n0, sizeof(double) * n0,
                                hipMemcpyHostToDevice, sp->stream.s));
  NC_CHECK_HIP(hipStreamSynchronize(sp->stream.s));

  nc_solve_opts ode = o.ode;
  if (ode.method == NC_METHOD_AUTO) ode.method = is_time_varying(model) ? NC_METHOD_RK45 : NC_METHOD_KRYLOV;

  nc_solve_stats tot{};
  DropWork work;
  DevBuf<int64_t> ids;
  std::vector<DevBuf<double>> kept;
  double tnow = tstart;
  int64_t adaptations = 0;
  int save_idx = 0;  // first saveat point not yet passed
  NC_REQUIRE(P == 0 || !o.records, NC_ERR_ARG, "training records come from plain (P = 0) solves");
  LearnWork lw;
  AdaptCtx ctx;
  ctx.model = model;
  ctx.records = o.records;
  ctx.lw = &lw;
  const double fallback_dt = o.adapter.fallback_dt > 0.0 ? o.adapter.fallback_dt : 1e-3 * (tend - tstart);
  double last_adapt_t = -INFINITY;
  bool last_learned = false;
  int64_t n_learned = 0, n_fallback = 0, max_states = sp->n;
  double step_hint = 0.0;
  while (tnow < tend) {
    std::unique_ptr<nc_matrix_s> Aown;
    std::unique_ptr<nc_sensmatrix_s> SA;
    nc_matrix_s* A = nullptr;
    if (P) {
      SA.reset(sensmatrix_create(sp.get(), model, o.precision));
      A = SA->A;
    } else {
      Aown.reset(matrix_create(sp.get(), model, o.precision));
      A = Aown.get();
    }
    std::unique_ptr<nc_solver_s> S(solver_create(A, &ode, SA.get()));
    solver_set_state_device(S.get(), tnow, ucur.p);
    // a restart continues with the step size the previous space's integrator reached
    if (ode.h0 <= 0.0) solver_seed_step(S.get(), step_hint);
    output_add_space(out, sp.get());
    EventSpec ev{o.fsptol, tend, false};
    auto advance = [&](double target, bool one_step) {
      if (S->method == NC_METHOD_KRYLOV)
        krylov_advance(S.get(), target, &ev, one_step);
      else
        rk_advance(S.get(), target, &ev, one_step);
    };
    if (n_save > 0) {
      // saveat semantics of DifferentialEquations: the start is saved when it is a
      // saveat point, then every point in (tnow, tend] the integration passes
      for (int k = 0; k < n_save; ++k)
        if (saveat[k] == tnow) output_add_snap(out, tnow, S.get());
      while (save_idx < n_save && saveat[save_idx] <= tnow) ++save_idx;
      for (;;) {
        const bool to_point = save_idx < n_save && saveat[save_idx] <= tend;
        const double target = to_point ? saveat[save_idx] : tend;
        advance(target, false);
        if (ev.fired) break;
        if (to_point) {
          output_add_snap(out, S->t, S.get());
          ++save_idx;
          while (save_idx < n_save && saveat[save_idx] == saveat[save_idx - 1]) {
            output_add_snap(out, S->t, S.get());
            ++save_idx;
          }
        }
        if (S->t >= tend) break;
      }
    } else {
      // empty saveat: every integrator step is saved, the start included
      output_add_snap(out, tnow, S.get());
      for (;;) {
        advance(tend, true);
        output_add_snap(out, S->t, S.get());
        if (ev.fired || S->t >= tend) break;
      }
    }
    NC_CHECK_HIP(hipStreamSynchronize(S->s));
    accumulate_stats(tot, S.get());
    tnow = S->t;
    step_hint = solver_step_hint(S.get());
    if (ev.fired && tnow < tend) {
      NC_REQUIRE(o.max_adaptations <= 0 || adaptations < o.max_adaptations, NC_ERR_NUMERIC,
                 "adaptive FSP exceeded max_adaptations");
      ++adaptations;
      const int64_t n_old = A->n;
      // sinks of every block (block 0: sink probabilities, others: their derivatives)
      std::vector<double> sinks((size_t)B * nr);
      NC_CHECK_HIP(hipMemcpy2D(sinks.data(), sizeof(double) * nr, S->u.p + n_old, sizeof(double) * S->Npb,
                               sizeof(double) * nr, B, hipMemcpyDeviceToHost));
      // the records of the previous adaptation are labelled by the mass reached now
      if (o.records) records_label(o.records, sp.get(), S->u.p, sp->stream.s);
      std::vector<double> dsinks;
      if (selective_kind(o.adapter.kind) || o.records) dsinks = sink_derivative(A, tnow, S->u.p, S->s);
      std::vector<const double*> vecs;
      for (int b = 0; b < B; ++b) vecs.push_back(S->u.p + b * S->Npb);
      // an event soon after a learned adaptation means it added too little: adapt plainly
      ctx.force_plain =
          o.adapter.kind == NC_ADAPTER_LEARNED && last_learned && tnow - last_adapt_t < fallback_dt;
      ctx.learned = false;
      const int64_t n_kept = n_old - adapt_space(sp.get(), o.adapter, vecs, dsinks.data(), tnow, tend, o.fsptol,
                                                 work, ids, kept, &ctx);
      if (o.adapter.kind == NC_ADAPTER_LEARNED) ++(ctx.learned ? n_learned : n_fallback);
      last_learned = ctx.learned;
      last_adapt_t = tnow;
      max_states = std::max<int64_t>(max_states, sp->n);
      if (o.verbose) print_update(tnow, sp->n);
      // numerical trick of fspsolve.jl:178-181: restart strictly below the budget
      double ssum = 0.0;
      for (int r = 0; r < nr; ++r) ssum += sinks[r];
      if (ssum >= tnow * o.fsptol / tend)
        for (int r = 0; r < nr; ++r) sinks[r] -= 2.220446049250313e-16;
      S.reset();
      SA.reset();
      Aown.reset();
      // unow = [p; zeros(new states); sinks] per block
      const int64_t n_new = sp->n;
      Nb = n_new + nr;
      ucur.alloc((size_t)(B * Nb));
      NC_CHECK_HIP(hipMemsetAsync(ucur.p, 0, sizeof(double) * B * Nb, sp->stream.s));
      for (int b = 0; b < B; ++b) {
        if (n_kept)
          NC_CHECK_HIP(hipMemcpyAsync(ucur.p + b * Nb, kept[b].p, sizeof(double) * n_kept, hipMemcpyDeviceToDevice,
                                      sp->stream.s));
        NC_CHECK_HIP(hipMemcpyAsync(ucur.p + b * Nb + n_new, sinks.data() + (size_t)b * nr, sizeof(double) * nr,
                                    hipMemcpyHostToDevice, sp->stream.s));
      }
      NC_CHECK_HIP(hipStreamSynchronize(sp->stream.s));
    } else {
      // the final slice is added once more (fspsolve.jl:189-194)
      output_add_snap(out, tnow, S.get());
      NC_CHECK_HIP(hipStreamSynchronize(S->s));
      if (o.records) records_label(o.records, sp.get(), S->u.p, sp->stream.s);
    }
  }
  tot.adaptations = adaptations;
  tot.learned_adaptations = n_learned;
  tot.fallback_adaptations = n_fallback;
  tot.max_states = max_states;
  if (st_out) *st_out = tot;
}

}  // namespace

extern "C" {

int nc_adapter_init(nc_space sp, const nc_adapter* ad) {
  return guard([&] {
    NC_REQUIRE(sp, NC_ERR_ARG, "null space");
    check_adapter(ad);
    space_expand_list(sp, ad->initial_step_count, all_reactions(sp->nr));
    NC_CHECK_HIP(hipStreamSynchronize(sp->stream.s));
  });
}

int nc_adapter_adapt(nc_space sp, const nc_adapter* ad, int64_t n, const double* p, const double* dsinks, double t,
                     double tend, double fsptol, int64_t* n_drop, int64_t* drop_ids) {
  return guard([&] {
    NC_REQUIRE(sp && n_drop, NC_ERR_ARG, "null argument");
    check_adapter(ad);
    NC_REQUIRE(ad->kind != NC_ADAPTER_LEARNED, NC_ERR_ARG,
               "the learned adapter's features need the model: it adapts inside nc_solve_adaptive");
    NC_REQUIRE(n == sp->n, NC_ERR_ARG, "p must hold one probability per state");
    NC_REQUIRE(n == 0 || p, NC_ERR_ARG, "null p");
    if (ad->kind == NC_ADAPTER_SELECTIVE_RSTEP) {
      NC_REQUIRE(n > 0, NC_ERR_ARG, "Empty `p` input in `adapt!`.");
      NC_REQUIRE(dsinks, NC_ERR_ARG, "the selective adapter needs the sink derivatives");
    }
    NC_REQUIRE(!ad->dropstates || n == 0 || drop_ids, NC_ERR_ARG, "null drop_ids");
    DevBuf<double> pd((size_t)std::max<int64_t>(n, 1));
    if (n) NC_CHECK_HIP(hipMemcpyAsync(pd.p, p, sizeof(double) * n, hipMemcpyHostToDevice, sp->stream.s));
    DropWork w;
    DevBuf<int64_t> ids;
    std::vector<DevBuf<double>> kept;
    const int64_t m = adapt_space(sp, *ad, {pd.p}, dsinks, t, tend, fsptol, w, ids, kept);
    if (m) NC_CHECK_HIP(hipMemcpy(drop_ids, ids.p, sizeof(int64_t) * m, hipMemcpyDeviceToHost));
    *n_drop = m;
  });
}

int nc_adaptive_opts_default(nc_adaptive_opts* o) {
  if (!o) return NC_ERR_ARG;
  std::memset(o, 0, sizeof(*o));
  o->adapter.kind = NC_ADAPTER_RSTEP;
  o->adapter.initial_step_count = 0;
  o->adapter.max_step_count = 1;
  o->adapter.dropstates = 0;
  o->precision = NC_FP64;
  o->max_adaptations = 100000;
  o->adapter.mlp = nullptr;
  o->adapter.threshold = 0.5;
  o->adapter.fallback_dt = 0.0;
  o->verbose = 0;
  o->records = nullptr;
  o->fsptol = 1e-6;
  return nc_solve_opts_default(&o->ode);
}

int nc_solve_adaptive(nc_model model, int64_t n0, const int32_t* states0, const double* p0, double t0, double t1,
                      int32_t n_save, const double* saveat, const nc_adaptive_opts* opts, nc_output* out,
                      nc_solve_stats* st) {
  return guard([&] {
    NC_REQUIRE(model && out, NC_ERR_ARG, "null argument");
    nc_adaptive_opts o{};
    if (opts)
      o = *opts;
    else
      nc_adaptive_opts_default(&o);
    auto O = std::make_unique<nc_output_s>();
    solve_adaptive_impl(model, 0, n0, states0, p0, nullptr, t0, t1, n_save, saveat, o, O.get(), st);
    *out = O.release();
  });
}

int nc_solve_sens_adaptive(nc_model model, int64_t n0, const int32_t* states0, const double* p0, const double* S0,
                           double t0, double t1, int32_t n_save, const double* saveat, const nc_adaptive_opts* opts,
                           nc_output* out, nc_solve_stats* st) {
  return guard([&] {
    NC_REQUIRE(model && out, NC_ERR_ARG, "null argument");
    NC_REQUIRE(model->has_grad, NC_ERR_ARG, "the model has no propensity gradients (CmeModelWithSensitivity)");
    NC_REQUIRE(model->np >= 1, NC_ERR_ARG, "the model has no parameters");
    nc_adaptive_opts o{};
    if (opts)
      o = *opts;
    else
      nc_adaptive_opts_default(&o);
    auto O = std::make_unique<nc_output_s>();
    solve_adaptive_impl(model, model->np, n0, states0, p0, S0, t0, t1, n_save, saveat, o, O.get(), st);
    *out = O.release();
  });
}

int nc_solve_adaptive_batch(int32_t n_inst, const nc_model* models, const int64_t* n0, const int32_t* states0,
                            const double* p0, double t0, double t1, int32_t n_save, const double* saveat,
                            const nc_adaptive_opts* opts, int32_t n_workers, nc_output* outs, nc_solve_stats* stats) {
  return guard([&] {
    NC_REQUIRE(n_inst >= 0, NC_ERR_ARG, "negative instance count");
    if (n_inst == 0) return;
    NC_REQUIRE(models && n0 && states0 && p0 && outs, NC_ERR_ARG, "null argument");
    nc_adaptive_opts o{};
    if (opts)
      o = *opts;
    else
      nc_adaptive_opts_default(&o);
    NC_REQUIRE(!o.records, NC_ERR_ARG, "training records come from single solves (nc_solve_adaptive)");
    // The workers launch eagerly: a stream capture in one thread is invalidated by the
    // allocations and synchronisations of the others (HIP does not keep thread-local
    // captures apart), so no worker captures graphs.
    o.ode.no_graphs = 1;
    // instance i's initial states and probabilities start at these offsets
    std::vector<int64_t> st_off(n_inst + 1, 0), p_off(n_inst + 1, 0);
    for (int32_t i = 0; i < n_inst; ++i) {
      NC_REQUIRE(models[i], NC_ERR_ARG, "null model of instance " + std::to_string(i));
      NC_REQUIRE(n0[i] >= 1, NC_ERR_ARG, "empty initial distribution of instance " + std::to_string(i));
      p_off[i + 1] = p_off[i] + n0[i];
      st_off[i + 1] = st_off[i] + n0[i] * models[i]->ns;
    }
    int dev = 0;
    NC_CHECK_HIP(hipGetDevice(&dev));
    const int W = std::max(1, std::min<int>(n_inst, n_workers > 0 ? n_workers : kBatchWorkers));
    std::vector<std::unique_ptr<nc_output_s>> res(n_inst);
    std::vector<nc_solve_stats> st(n_inst);
    std::atomic<int32_t> next{0};
    std::atomic<bool> failed{false};
    std::mutex err_mu;
    int err_code = NC_OK;
    std::string err_msg;
    std::vector<int32_t> deferred;  // ran out of device memory next to the others: run alone
    auto fail = [&](int32_t i, int code, const char* what) {
      std::lock_guard<std::mutex> lk(err_mu);
      if (err_code == NC_OK) {
        err_code = code;
        err_msg = "instance " + std::to_string(i) + ": " + what;
      }
      failed = true;
    };
    // Each worker runs whole solves, one instance at a time, on streams of its own; every
    // object a solve creates is created, used and freed on its worker.
    auto solve_one = [&](int32_t i) {
      auto O = std::make_unique<nc_output_s>();
      solve_adaptive_impl(models[i], 0, n0[i], states0 + st_off[i], p0 + p_off[i], nullptr, t0, t1, n_save, saveat, o,
                          O.get(), &st[i]);
      res[i] = std::move(O);
    };
    auto work = [&] {
      if (hipSetDevice(dev) != hipSuccess) {
        fail(-1, NC_ERR_HIP, "hipSetDevice failed on a worker thread");
        return;
      }
      PrivateStreams scope;
      for (;;) {
        const int32_t i = next.fetch_add(1);
        if (i >= n_inst || failed) break;
        try {
          solve_one(i);
        } catch (const Error& e) {
          if (e.code == NC_ERR_ALLOC && W > 1) {
            std::lock_guard<std::mutex> lk(err_mu);
            deferred.push_back(i);
          } else {
            fail(i, e.code, e.what());
          }
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
    if (failed) throw Error(err_code, err_msg);  // the finished outputs are freed with res
    // instances whose device memory ran out while others ran: again, one at a time, with
    // the cached blocks returned to the driver first
    std::sort(deferred.begin(), deferred.end());
    for (int32_t i : deferred) {
      pool_trim();
      try {
        solve_one(i);
      } catch (const Error& e) {
        throw Error(e.code, "instance " + std::to_string(i) + ": " + e.what());
      }
    }
    for (int32_t i = 0; i < n_inst; ++i) {
      outs[i] = res[i].release();
      if (stats) stats[i] = st[i];
    }
  });
}

int nc_output_get_sens(nc_output o, int32_t k, int32_t param, double* S, double* dsinks) {
  return guard([&] {
    NC_REQUIRE(o && k >= 0 && k < (int32_t)o->snaps.size(), NC_ERR_ARG, "snapshot index out of range");
    NC_REQUIRE(param >= 0 && param < o->P, NC_ERR_ARG, "parameter index out of range");
    const auto& P = *o->snaps[k];
    const int64_t n = o->spaces[P.space]->n, Nb = n + o->nr;
    const double* base = P.u.p + (size_t)(param + 1) * Nb;
    if (S && n) NC_CHECK_HIP(hipMemcpy(S, base, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (dsinks) NC_CHECK_HIP(hipMemcpy(dsinks, base + n, sizeof(double) * o->nr, hipMemcpyDeviceToHost));
  });
}

int nc_output_get_fspvec(nc_output o, int32_t k, int32_t param, nc_fspvec* p, double* sinks) {
  return guard([&] {
    NC_REQUIRE(o && p && k >= 0 && k < (int32_t)o->snaps.size(), NC_ERR_ARG, "snapshot index out of range");
    NC_REQUIRE(param >= -1 && param < o->P, NC_ERR_ARG, "parameter index out of range");
    const auto& P = *o->snaps[k];
    const auto& Sp = *o->spaces[P.space];
    const int64_t Nb = Sp.n + o->nr;
    const double* base = P.u.p + (size_t)(param + 1) * Nb;
    Stream st;
    *p = fspvec_from_device(o->ns, Sp.n, Sp.states, base, st.s);
    if (sinks) NC_CHECK_HIP(hipMemcpy(sinks, base + Sp.n, sizeof(double) * o->nr, hipMemcpyDeviceToHost));
  });
}

int nc_output_destroy(nc_output o) {
  return guard([&] { delete o; });
}

int nc_output_count(nc_output o, int32_t* n_snapshots, int32_t* n_spaces) {
  return guard([&] {
    NC_REQUIRE(o, NC_ERR_ARG, "null output");
    if (n_snapshots) *n_snapshots = (int32_t)o->snaps.size();
    if (n_spaces) *n_spaces = (int32_t)o->spaces.size();
  });
}

int nc_output_info(nc_output o, int32_t k, double* t, int64_t* n_states, int32_t* space_id) {
  return guard([&] {
    NC_REQUIRE(o && k >= 0 && k < (int32_t)o->snaps.size(), NC_ERR_ARG, "snapshot index out of range");
    const auto& P = *o->snaps[k];
    if (t) *t = P.t;
    if (n_states) *n_states = o->spaces[P.space]->n;
    if (space_id) *space_id = P.space;
  });
}

int nc_output_get(nc_output o, int32_t k, int32_t* states, double* p, double* sinks) {
  return guard([&] {
    NC_REQUIRE(o && k >= 0 && k < (int32_t)o->snaps.size(), NC_ERR_ARG, "snapshot index out of range");
    const auto& P = *o->snaps[k];
    const auto& Sp = *o->spaces[P.space];
    if (states && Sp.n)
      NC_CHECK_HIP(hipMemcpy(states, Sp.states->p, sizeof(int32_t) * Sp.n * o->ns, hipMemcpyDeviceToHost));
    if (p && Sp.n) NC_CHECK_HIP(hipMemcpy(p, P.u.p, sizeof(double) * Sp.n, hipMemcpyDeviceToHost));
    if (sinks) NC_CHECK_HIP(hipMemcpy(sinks, P.u.p + Sp.n, sizeof(double) * o->nr, hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
