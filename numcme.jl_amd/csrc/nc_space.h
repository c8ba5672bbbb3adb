// Device-resident sparse FSP state space (internal definitions shared by the space,
// matrix and adapter sources).
//
// Layout in HBM (n states, ns species, nr reactions):
//   states      int32  [cap][ns]   row-major, index order == reference insertion order
//   state_conn  int32  [cap][nr]   index of x_i - s_r, or -1
//   sink_mask   uint32 [cap]       bit r set when x_i + s_r >= 0 and is not in the space
//   table       uint32 [2^k]       open-addressing hash (linear probing): state index,
//                                  HT_EMPTY, or HT_CAND|candidate-position while a level
//                                  of expansion is being deduplicated
#pragma once

#include "nc_internal.h"

namespace nc {

constexpr int kMaxSpecies = 16;
constexpr int kMaxReactions = 32;
constexpr uint32_t HT_EMPTY = 0xFFFFFFFFu;
constexpr uint32_t HT_CAND = 0x80000000u;

__host__ __device__ inline uint64_t mix64(uint64_t z) {
  // splitmix64 finaliser (Steele, Lea & Flood 2014)
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__host__ __device__ inline uint64_t hash_state(const int32_t* x, int ns) {
  uint64_t h = 0x9E3779B97F4A7C15ull * (uint64_t)(ns + 1);
  for (int s = 0; s < ns; ++s) h = mix64(h ^ (uint64_t)(uint32_t)x[s]);
  return h;
}

__device__ inline bool state_eq(const int32_t* a, const int32_t* b, int ns) {
  for (int s = 0; s < ns; ++s)
    if (a[s] != b[s]) return false;
  return true;
}

// Index of state x, or -1. Valid when the table holds no candidate tags.
__device__ inline int32_t ht_find(const uint32_t* table, uint64_t mask, const int32_t* states, int ns,
                                  const int32_t* x) {
  uint64_t h = hash_state(x, ns) & mask;
  for (;;) {
    uint32_t v = table[h];
    if (v == HT_EMPTY) return -1;
    if (!(v & HT_CAND) && state_eq(states + (int64_t)v * ns, x, ns)) return (int32_t)v;
    h = (h + 1) & mask;
  }
}

}  // namespace nc

struct nc_space_s {
  int ns = 0, nr = 0;
  std::vector<int32_t> stoich;  // ns x nr, column-major (host copy)
  nc::Stream stream;
  int64_t n = 0;    // number of states
  int64_t cap = 0;  // allocated rows
  nc::DevBuf<int32_t> d_stoich;  // ns x nr column-major
  nc::DevBuf<int32_t> states;
  nc::DevBuf<int32_t> state_conn;
  nc::DevBuf<uint32_t> sink_mask;
  nc::DevBuf<uint32_t> table;
  uint64_t table_cap = 0;
  // scratch reused across calls
  nc::DevBuf<int32_t> cands, slot_of, explorables, remap;
  nc::DevBuf<uint32_t> flags, ranks;
  nc::DevBuf<unsigned char> scan_tmp;
  uint64_t version = 0;  // bumped on every change of the state list
  uint64_t uid = 0;      // unique per space in the process
};

namespace nc {
// StateSpaceSparse(stoich, init_states): duplicates and negative states are dropped
nc_space_s* space_create(int ns, int nr, const int32_t* stoich, int64_t n_init, const int32_t* init_states);
void space_reserve(nc_space_s* sp, int64_t rows);
void space_rebuild_table(nc_space_s* sp, int64_t min_entries);
void space_add_candidates(nc_space_s* sp, int64_t n_cand);  // candidates already in sp->cands
// expand! over the reactions in rlist (0-based; callers pass every reaction for "all").
// With `select` (device flags, one per state) the first level starts only from the
// explorable states whose flag is nonzero; later levels explore every state added.
void space_expand_list(nc_space_s* sp, int levels, const std::vector<int32_t>& rlist,
                       const uint32_t* select = nullptr);
// deleteat! with sorted unique device ids; compacts nvec companion vectors alongside
void space_delete_device(nc_space_s* sp, const int64_t* d_ids, int64_t m, int nvec, const double* const* vin,
                         double* const* vout);
}  // namespace nc
