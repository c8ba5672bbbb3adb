// Device-resident sparse FSP state space.
//
// Replaces StateSpaceSparse (NumCME.jl src/statespace/sparse/sparsestatespace.jl) with a
// GPU open-addressing hash table. The reference's sequential semantics are kept exactly:
//
//  * expand! (sparsestatespace.jl:153-194) explores the explorable states last-first (its
//    Deque pop! takes from the back) and, within one state, reactions in list order. Here
//    every candidate gets its position in that sequential order; candidates are inserted
//    in parallel and a key's slot keeps the smallest position (atomicMin), so the first
//    occurrence wins as in _addstates! (:217-226). A prefix sum over the winners assigns
//    the reference's indices.
//  * The connectivity fill (:241-266) writes values that depend only on the final set of
//    states, so it runs one thread per (new state, reaction).
//  * deleteat! (:276-331) is a keep-flag scan, a compaction with index remapping, a hash
//    rebuild and a sink recomputation.
#include <atomic>

#include "nc_space.h"

namespace nc {

namespace {

struct RList {
  int32_t n;
  int32_t r[kMaxReactions];
};

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_table_insert_all(uint32_t* table, uint64_t mask, const int32_t* states, int ns, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t h = hash_state(states + i * ns, ns) & mask;
  for (;;) {
    uint32_t old = atomicCAS(&table[h], HT_EMPTY, (uint32_t)i);
    if (old == HT_EMPTY) return;
    h = (h + 1) & mask;
  }
}

// Explorables of the first level: states with a sink on one of the listed reactions (and,
// when `select` is given, a nonzero select flag).
__global__ void k_flag_explorable(const uint32_t* sink_mask, int64_t n, uint32_t rmask, const uint32_t* select,
                                  uint32_t* flags) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) flags[i] = ((sink_mask[i] & rmask) && (!select || select[i])) ? 1u : 0u;
}

__global__ void k_scatter_flagged(const uint32_t* flags, const uint32_t* ranks, int64_t n, int32_t* out) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n && flags[i]) out[ranks[i]] = (int32_t)i;
}

// Candidate `pos` in the reference's order: explorable E[nE-1-pos/|R'|] (LIFO), reaction
// R'[pos % |R'|]. `explorables == nullptr` means E[e] = base + e.
__global__ void k_gen_candidates(const int32_t* states, const int32_t* stoich, int ns, const int32_t* explorables,
                                 int64_t base, int64_t nE, RList rl, int32_t* cands) {
  int64_t pos = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (pos >= nE * rl.n) return;
  int64_t e = nE - 1 - pos / rl.n;
  int r = rl.r[pos % rl.n];
  int64_t src = explorables ? explorables[e] : base + e;
  for (int s = 0; s < ns; ++s) cands[pos * ns + s] = states[src * ns + s] + stoich[s + r * ns];
}

__global__ void k_cand_insert(uint32_t* table, uint64_t mask, const int32_t* states, const int32_t* cands, int ns,
                              int64_t n_cand, int32_t* slot_of) {
  int64_t pos = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (pos >= n_cand) return;
  int32_t x[kMaxSpecies];
  bool ok = true;
  for (int s = 0; s < ns; ++s) {
    x[s] = cands[pos * ns + s];
    ok = ok && (x[s] >= 0);
  }
  if (!ok) {  // _is_all_nonnegative (sparsestatespace.jl:197-202)
    slot_of[pos] = -1;
    return;
  }
  const uint32_t tag = HT_CAND | (uint32_t)pos;
  uint64_t h = hash_state(x, ns) & mask;
  for (;;) {
    uint32_t v = __hip_atomic_load(&table[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (v == HT_EMPTY) {
      uint32_t old = atomicCAS(&table[h], HT_EMPTY, tag);
      if (old == HT_EMPTY) {
        slot_of[pos] = (int32_t)h;
        return;
      }
      v = old;
    }
    if (v & HT_CAND) {
      // every tag ever stored in this slot belongs to the same key, so comparing against
      // the one read is enough even if a smaller position replaces it concurrently
      if (state_eq(cands + (int64_t)(v & ~HT_CAND) * ns, x, ns)) {
        atomicMin(&table[h], tag);
        slot_of[pos] = (int32_t)h;
        return;
      }
    } else if (state_eq(states + (int64_t)v * ns, x, ns)) {
      slot_of[pos] = -1;  // already in the space
      return;
    }
    h = (h + 1) & mask;
  }
}

__global__ void k_first_flags(const uint32_t* table, const int32_t* slot_of, int64_t n_cand, uint32_t* flags) {
  int64_t pos = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (pos >= n_cand) return;
  int32_t sl = slot_of[pos];
  flags[pos] = (sl >= 0 && table[sl] == (HT_CAND | (uint32_t)pos)) ? 1u : 0u;
}

__global__ void k_commit(uint32_t* table, const int32_t* slot_of, const uint32_t* flags, const uint32_t* ranks,
                         const int32_t* cands, int ns, int nr, int64_t n_cand, int64_t n_old, int32_t* states,
                         int32_t* state_conn, uint32_t* sink_mask) {
  int64_t pos = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (pos >= n_cand || !flags[pos]) return;
  int64_t idx = n_old + ranks[pos];
  for (int s = 0; s < ns; ++s) states[idx * ns + s] = cands[pos * ns + s];
  for (int r = 0; r < nr; ++r) state_conn[idx * nr + r] = -1;
  sink_mask[idx] = 0u;
  table[slot_of[pos]] = (uint32_t)idx;
}

// Connectivity of the states [n_old, n_new): sparsestatespace.jl:241-266.
__global__ void k_connect_new(const uint32_t* table, uint64_t mask, const int32_t* states, const int32_t* stoich,
                              int ns, int nr, int64_t n_old, int64_t n_new, int32_t* state_conn,
                              uint32_t* sink_mask) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (n_new - n_old) * nr) return;
  int64_t idx = n_old + t / nr;
  int r = (int)(t % nr);
  int32_t y[kMaxSpecies], z[kMaxSpecies];
  bool yok = true, zok = true;
  for (int s = 0; s < ns; ++s) {
    int32_t xs = states[idx * ns + s], d = stoich[s + r * ns];
    y[s] = xs - d;
    z[s] = xs + d;
    yok = yok && y[s] >= 0;
    zok = zok && z[s] >= 0;
  }
  if (yok) {
    int32_t j = ht_find(table, mask, states, ns, y);
    if (j >= 0) {
      state_conn[idx * nr + r] = j;
      atomicAnd(&sink_mask[j], ~(1u << r));
    }
  }
  if (zok) {
    int32_t k = ht_find(table, mask, states, ns, z);
    if (k < 0)
      atomicOr(&sink_mask[idx], 1u << r);
    else
      state_conn[(int64_t)k * nr + r] = (int32_t)idx;
  }
}

__global__ void k_mark_ids(const int64_t* ids, int64_t m, uint32_t* keep) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < m) keep[ids[i]] = 0u;
}

__global__ void k_compact(const uint32_t* keep, const uint32_t* ranks, int64_t n, int ns, int nr, const int32_t* states,
                          const int32_t* state_conn, const uint32_t* sink_mask, int32_t* states2, int32_t* conn2,
                          uint32_t* sink2) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n || !keep[i]) return;
  int64_t d = ranks[i];
  for (int s = 0; s < ns; ++s) states2[d * ns + s] = states[i * ns + s];
  for (int r = 0; r < nr; ++r) {
    int32_t c = state_conn[i * nr + r];
    conn2[d * nr + r] = (c >= 0 && keep[c]) ? (int32_t)ranks[c] : -1;
  }
  sink2[d] = sink_mask[i];
}

// sparsestatespace.jl:318-329: new sinks for states whose successor was deleted.
__global__ void k_recompute_sinks(const uint32_t* table, uint64_t mask, const int32_t* states, const int32_t* stoich,
                                  int ns, int nr, int64_t n, uint32_t* sink_mask) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t m = sink_mask[i];
  int32_t z[kMaxSpecies];
  for (int r = 0; r < nr; ++r) {
    if (m & (1u << r)) continue;
    bool ok = true;
    for (int s = 0; s < ns; ++s) {
      z[s] = states[i * ns + s] + stoich[s + r * ns];
      ok = ok && z[s] >= 0;
    }
    if (ok && ht_find(table, mask, states, ns, z) < 0) m |= 1u << r;
  }
  sink_mask[i] = m;
}

__global__ void k_lookup(const uint32_t* table, uint64_t mask, const int32_t* states, int ns, const int32_t* q,
                         int64_t nq, int64_t* out) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nq) return;
  int32_t x[kMaxSpecies];
  bool ok = true;
  for (int s = 0; s < ns; ++s) {
    x[s] = q[i * ns + s];
    ok = ok && x[s] >= 0;
  }
  out[i] = (ok && mask) ? (int64_t)ht_find(table, mask, states, ns, x) : -1;
}

constexpr int kBlock = 256;

}  // namespace

void space_reserve(nc_space_s* sp, int64_t rows) {
  if (rows <= sp->cap) return;
  int64_t cap = sp->cap ? sp->cap : 1024;
  while (cap < rows) cap *= 2;
  sp->states.grow_keep((size_t)cap * sp->ns, (size_t)sp->n * sp->ns, sp->stream.s);
  sp->state_conn.grow_keep((size_t)cap * sp->nr, (size_t)sp->n * sp->nr, sp->stream.s);
  sp->sink_mask.grow_keep((size_t)cap, (size_t)sp->n, sp->stream.s);
  sp->cap = cap;
}

void space_rebuild_table(nc_space_s* sp, int64_t min_entries) {
  uint64_t cap = 1024;
  while (cap < 2 * (uint64_t)min_entries) cap *= 2;
  NC_REQUIRE(cap <= (1ull << 31), NC_ERR_OVERFLOW, "state space too large for the hash table");
  if (cap > sp->table_cap || !sp->table.p) {
    sp->table.alloc(cap);
    sp->table_cap = cap;
  }
  hipStream_t s = sp->stream.s;
  k_fill_u32<<<grid_for(sp->table_cap, kBlock), kBlock, 0, s>>>(sp->table.p, (int64_t)sp->table_cap, HT_EMPTY);
  if (sp->n)
    k_table_insert_all<<<grid_for(sp->n, kBlock), kBlock, 0, s>>>(sp->table.p, sp->table_cap - 1, sp->states.p,
                                                                   sp->ns, sp->n);
  NC_CHECK_HIP(hipGetLastError());
}

// Deduplicate and append the n_cand candidates held in sp->cands (candidate order is the
// reference's _addstates! order), then fill connectivity for the new states.
void space_add_candidates(nc_space_s* sp, int64_t n_cand) {
  if (n_cand <= 0) return;
  hipStream_t s = sp->stream.s;
  NC_REQUIRE(sp->n + n_cand < (int64_t)0x7FFFFFF0, NC_ERR_OVERFLOW, "state space exceeds 2^31 states");
  space_reserve(sp, sp->n + n_cand);
  if ((uint64_t)(2 * (sp->n + n_cand)) > sp->table_cap) space_rebuild_table(sp, sp->n + n_cand);
  sp->slot_of.reserve(n_cand);
  sp->flags.reserve(n_cand);
  sp->ranks.reserve(n_cand);
  const uint64_t mask = sp->table_cap - 1;
  k_cand_insert<<<grid_for(n_cand, kBlock), kBlock, 0, s>>>(sp->table.p, mask, sp->states.p, sp->cands.p, sp->ns,
                                                             n_cand, sp->slot_of.p);
  k_first_flags<<<grid_for(n_cand, kBlock), kBlock, 0, s>>>(sp->table.p, sp->slot_of.p, n_cand, sp->flags.p);
  NC_CHECK_HIP(hipGetLastError());
  uint32_t added = 0;
  exclusive_scan_u32(sp->flags.p, sp->ranks.p, n_cand, sp->scan_tmp, s, &added);
  const int64_t n_old = sp->n;
  k_commit<<<grid_for(n_cand, kBlock), kBlock, 0, s>>>(sp->table.p, sp->slot_of.p, sp->flags.p, sp->ranks.p,
                                                        sp->cands.p, sp->ns, sp->nr, n_cand, n_old, sp->states.p,
                                                        sp->state_conn.p, sp->sink_mask.p);
  NC_CHECK_HIP(hipGetLastError());
  sp->n = n_old + added;
  if (added) {
    int64_t work = (int64_t)added * sp->nr;
    k_connect_new<<<grid_for(work, kBlock), kBlock, 0, s>>>(sp->table.p, mask, sp->states.p, sp->d_stoich.p, sp->ns,
                                                             sp->nr, n_old, sp->n, sp->state_conn.p,
                                                             sp->sink_mask.p);
    NC_CHECK_HIP(hipGetLastError());
    sp->version++;
  }
}

static void space_expand(nc_space_s* sp, int levels, const std::vector<int32_t>& rlist,
                         const uint32_t* select = nullptr) {
  if (levels <= 0) return;  // sparsestatespace.jl:154-156
  hipStream_t s = sp->stream.s;
  RList rl{};
  rl.n = (int32_t)rlist.size();
  uint32_t rmask = 0;
  for (int k = 0; k < rl.n; ++k) {
    rl.r[k] = rlist[k];
    rmask |= 1u << rlist[k];
  }
  // first-level explorables: ascending indices with a sink on some listed reaction (:165-173)
  int64_t nE = 0;
  if (sp->n) {
    sp->flags.reserve(sp->n);
    sp->ranks.reserve(sp->n);
    sp->explorables.reserve(sp->n);
    k_flag_explorable<<<grid_for(sp->n, kBlock), kBlock, 0, s>>>(sp->sink_mask.p, sp->n, rmask, select,
                                                                   sp->flags.p);
    uint32_t cnt = 0;
    exclusive_scan_u32(sp->flags.p, sp->ranks.p, sp->n, sp->scan_tmp, s, &cnt);
    k_scatter_flagged<<<grid_for(sp->n, kBlock), kBlock, 0, s>>>(sp->flags.p, sp->ranks.p, sp->n, sp->explorables.p);
    NC_CHECK_HIP(hipGetLastError());
    nE = cnt;
  }
  const int32_t* E = sp->explorables.p;
  int64_t base = 0;
  for (int level = 0; level < levels; ++level) {  // :176-192
    int64_t n_cand = nE * rl.n;
    if (n_cand == 0) break;
    NC_REQUIRE(n_cand < (int64_t)0x7FFFFFF0, NC_ERR_OVERFLOW, "too many expansion candidates");
    sp->cands.reserve((size_t)n_cand * sp->ns);
    k_gen_candidates<<<grid_for(n_cand, kBlock), kBlock, 0, s>>>(sp->states.p, sp->d_stoich.p, sp->ns, E, base, nE,
                                                                  rl, sp->cands.p);
    NC_CHECK_HIP(hipGetLastError());
    const int64_t n_old = sp->n;
    space_add_candidates(sp, n_cand);
    // the next level explores the states just added, in ascending order (:189-191)
    E = nullptr;
    base = n_old;
    nE = sp->n - n_old;
  }
}

__global__ void k_compact_vec(const uint32_t* keep, const uint32_t* ranks, int64_t n, const double* in, double* out) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n && keep[i]) out[ranks[i]] = in[i];
}

// deleteat! with the (sorted, unique, in-range) ids already on the device. The vectors
// vin[k] (length n before the call) are compacted the same way into vout[k].
void space_delete_device(nc_space_s* sp, const int64_t* d_ids, int64_t m, int nvec, const double* const* vin,
                         double* const* vout) {
  hipStream_t s = sp->stream.s;
  if (sp->n == 0) return;
  const int64_t n = sp->n;
  sp->flags.reserve(n);
  sp->ranks.reserve(n);
  k_fill_u32<<<grid_for(n, kBlock), kBlock, 0, s>>>(sp->flags.p, n, 1u);
  if (m) k_mark_ids<<<grid_for(m, kBlock), kBlock, 0, s>>>(d_ids, m, sp->flags.p);
  uint32_t kept = 0;
  exclusive_scan_u32(sp->flags.p, sp->ranks.p, n, sp->scan_tmp, s, &kept);
  for (int k = 0; k < nvec; ++k)
    k_compact_vec<<<grid_for(n, kBlock), kBlock, 0, s>>>(sp->flags.p, sp->ranks.p, n, vin[k], vout[k]);
  DevBuf<int32_t> st2((size_t)sp->cap * sp->ns), cn2((size_t)sp->cap * sp->nr);
  DevBuf<uint32_t> sk2((size_t)sp->cap);
  k_compact<<<grid_for(n, kBlock), kBlock, 0, s>>>(sp->flags.p, sp->ranks.p, n, sp->ns, sp->nr, sp->states.p,
                                                    sp->state_conn.p, sp->sink_mask.p, st2.p, cn2.p, sk2.p);
  NC_CHECK_HIP(hipGetLastError());
  NC_CHECK_HIP(hipStreamSynchronize(s));
  sp->states = std::move(st2);
  sp->state_conn = std::move(cn2);
  sp->sink_mask = std::move(sk2);
  sp->n = kept;
  sp->version++;
  if (sp->n == 0) return;  // :306-308
  space_rebuild_table(sp, sp->n);
  k_recompute_sinks<<<grid_for(sp->n, kBlock), kBlock, 0, s>>>(sp->table.p, sp->table_cap - 1, sp->states.p,
                                                                sp->d_stoich.p, sp->ns, sp->nr, sp->n,
                                                                sp->sink_mask.p);
  NC_CHECK_HIP(hipGetLastError());
}

static void space_delete(nc_space_s* sp, int64_t m, const int64_t* ids_host) {
  for (int64_t k = 0; k < m; ++k) {
    NC_REQUIRE(ids_host[k] >= 0 && ids_host[k] < sp->n, NC_ERR_ARG, "deleteat!: index out of range");
    NC_REQUIRE(k == 0 || ids_host[k] > ids_host[k - 1], NC_ERR_ARG, "deleteat!: indices must be unique and sorted");
  }
  if (sp->n == 0) return;
  DevBuf<int64_t> d_ids(m ? m : 1);
  if (m) NC_CHECK_HIP(hipMemcpyAsync(d_ids.p, ids_host, m * sizeof(int64_t), hipMemcpyHostToDevice, sp->stream.s));
  space_delete_device(sp, d_ids.p, m, 0, nullptr, nullptr);
}

nc_space_s* space_create(int ns, int nr, const int32_t* stoich, int64_t n_init, const int32_t* init_states) {
  NC_REQUIRE(ns >= 1 && ns <= kMaxSpecies, NC_ERR_UNSUPPORTED, "species count must be in 1..16");
  NC_REQUIRE(nr >= 1 && nr <= kMaxReactions, NC_ERR_UNSUPPORTED, "reaction count must be in 1..32");
  NC_REQUIRE(stoich, NC_ERR_ARG, "null stoichiometry matrix");
  NC_REQUIRE(n_init >= 0 && (n_init == 0 || init_states), NC_ERR_ARG, "bad initial state list");
  int dev_count = 0;
  if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0)
    throw Error(NC_ERR_NO_DEVICE, "no HIP device available");
  auto sp = new nc_space_s();
  static std::atomic<uint64_t> next_uid{1};
  sp->uid = next_uid++;
  try {
    sp->ns = ns;
    sp->nr = nr;
    sp->stoich.assign(stoich, stoich + (size_t)ns * nr);
    sp->d_stoich.alloc((size_t)ns * nr);
    // on the space's stream from its own host copy (the constructor ends with a
    // synchronisation of that stream)
    NC_CHECK_HIP(hipMemcpyAsync(sp->d_stoich.p, sp->stoich.data(), sizeof(int32_t) * ns * nr, hipMemcpyHostToDevice,
                                sp->stream.s));
    space_reserve(sp, std::max<int64_t>(n_init, 1));
    space_rebuild_table(sp, std::max<int64_t>(n_init, 1));
    if (n_init) {
      sp->cands.reserve((size_t)n_init * ns);
      NC_CHECK_HIP(hipMemcpyAsync(sp->cands.p, init_states, sizeof(int32_t) * ns * n_init, hipMemcpyHostToDevice,
                                  sp->stream.s));
      space_add_candidates(sp, n_init);  // _addstates!(space, initstates): :122
    }
    NC_CHECK_HIP(hipStreamSynchronize(sp->stream.s));
  } catch (...) {
    delete sp;
    throw;
  }
  return sp;
}

void space_expand_list(nc_space_s* sp, int levels, const std::vector<int32_t>& rlist, const uint32_t* select) {
  space_expand(sp, levels, rlist, select);
}

}  // namespace nc

using namespace nc;

extern "C" {

int nc_space_create(int32_t ns, int32_t nr, const int32_t* stoich, int64_t n_init, const int32_t* init_states,
                    nc_space* out) {
  return guard([&] {
    NC_REQUIRE(out, NC_ERR_ARG, "null output handle");
    *out = space_create(ns, nr, stoich, n_init, init_states);
  });
}

int nc_space_destroy(nc_space sp) {
  return guard([&] { delete sp; });
}

int nc_space_expand(nc_space sp, int32_t levels, int32_t n_only, const int32_t* only) {
  return guard([&] {
    NC_REQUIRE(sp, NC_ERR_ARG, "null space");
    std::vector<int32_t> rl;
    if (n_only > 0) {
      NC_REQUIRE(only, NC_ERR_ARG, "null reaction list");
      NC_REQUIRE(n_only <= kMaxReactions, NC_ERR_ARG, "reaction list too long");
      for (int k = 0; k < n_only; ++k) {
        NC_REQUIRE(only[k] >= 0 && only[k] < sp->nr, NC_ERR_ARG, "reaction index out of range");
        rl.push_back(only[k]);
      }
    } else {
      for (int r = 0; r < sp->nr; ++r) rl.push_back(r);  // isempty(onlyreactions) -> all (:163)
    }
    space_expand(sp, levels, rl);
    NC_CHECK_HIP(hipStreamSynchronize(sp->stream.s));
  });
}

int nc_space_delete(nc_space sp, int64_t m, const int64_t* ids) {
  return guard([&] {
    NC_REQUIRE(sp, NC_ERR_ARG, "null space");
    NC_REQUIRE(m >= 0 && (m == 0 || ids), NC_ERR_ARG, "bad id list");
    space_delete(sp, m, ids);
    NC_CHECK_HIP(hipStreamSynchronize(sp->stream.s));
  });
}

int nc_space_count(nc_space sp, int64_t* n, int32_t* nsinks) {
  return guard([&] {
    NC_REQUIRE(sp, NC_ERR_ARG, "null space");
    if (n) *n = sp->n;
    if (nsinks) *nsinks = sp->nr;
  });
}

int nc_space_get_states(nc_space sp, int32_t* out) {
  return guard([&] {
    NC_REQUIRE(sp && (out || sp->n == 0), NC_ERR_ARG, "null argument");
    if (sp->n)
      NC_CHECK_HIP(hipMemcpy(out, sp->states.p, sizeof(int32_t) * sp->n * sp->ns, hipMemcpyDeviceToHost));
  });
}

int nc_space_get_state_conn(nc_space sp, int32_t* out) {
  return guard([&] {
    NC_REQUIRE(sp && (out || sp->n == 0), NC_ERR_ARG, "null argument");
    if (sp->n)
      NC_CHECK_HIP(hipMemcpy(out, sp->state_conn.p, sizeof(int32_t) * sp->n * sp->nr, hipMemcpyDeviceToHost));
  });
}

int nc_space_get_sink_conn(nc_space sp, int32_t* out) {
  return guard([&] {
    NC_REQUIRE(sp && (out || sp->n == 0), NC_ERR_ARG, "null argument");
    if (!sp->n) return;
    std::vector<uint32_t> m(sp->n);
    NC_CHECK_HIP(hipMemcpy(m.data(), sp->sink_mask.p, sizeof(uint32_t) * sp->n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < sp->n; ++i)
      for (int r = 0; r < sp->nr; ++r) out[i * sp->nr + r] = (m[i] >> r) & 1u ? r : -1;
  });
}

int nc_space_lookup(nc_space sp, int64_t n, const int32_t* q, int64_t* out) {
  return guard([&] {
    NC_REQUIRE(sp && n >= 0 && (n == 0 || (q && out)), NC_ERR_ARG, "null argument");
    if (n == 0) return;
    hipStream_t s = sp->stream.s;
    DevBuf<int32_t> dq((size_t)n * sp->ns);
    DevBuf<int64_t> dout((size_t)n);
    NC_CHECK_HIP(hipMemcpyAsync(dq.p, q, sizeof(int32_t) * n * sp->ns, hipMemcpyHostToDevice, s));
    k_lookup<<<grid_for(n, kBlock), kBlock, 0, s>>>(sp->table.p, sp->n ? sp->table_cap - 1 : 0, sp->states.p, sp->ns,
                                                     dq.p, n, dout.p);
    NC_CHECK_HIP(hipGetLastError());
    NC_CHECK_HIP(hipMemcpyAsync(out, dout.p, sizeof(int64_t) * n, hipMemcpyDeviceToHost, s));
    NC_CHECK_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
