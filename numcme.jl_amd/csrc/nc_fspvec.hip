// Device-resident sparse FSP vectors: construction, download, total, marginals
// (sum over species) and dense conversion — FspVectorSparse of NumCME.jl
// (src/fspvector/fspvector.jl:13-129), kept on the GPU so that snapshots need no host
// round trip and no per-snapshot host dictionary.
//
// Marginal sum(p, dims) (fspvector.jl:66-99). The reference walks the states in order and
// gives each new reduced state the next index, so the result lists the reduced states in
// order of first appearance, and each value is a left fold of the values in state order.
// Here:
//   1. pack each state's kept coordinates (offset by their minimum) into a 64-bit key;
//   2. stable radix sort of (key, index): equal keys keep their state order;
//   3. segment heads -> group of every entry; the first index of a group is its head's;
//   4. sort the groups by their first index: the reference's order of first appearance;
//   5. one thread per group folds its values left to right in state order.
// Step 5 adds in the reference's order, so the values are bit-identical to its loop.
//
// Array(p) (fspvector.jl:107-129): bounds = max count + 1 per species; the dense array is
// column-major (species 0 fastest, as Julia's Array). For repeated states the last one
// wins, as the reference's assignment loop: each cell keeps the largest state index that
// maps to it (atomicMax), then that entry's value is written.
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>

#include "nc_fspvec.h"
#include "nc_internal.h"

using namespace nc;

namespace nc {

namespace {

constexpr int kB = 256;
constexpr int kMaxNs = 16;

__global__ void k_minmax(int64_t n, int ns, const int32_t* states, int32_t* mn, int32_t* mx) {
  // mn, mx: ns entries each, pre-set to INT_MAX / INT_MIN
  __shared__ int32_t smn[kMaxNs], smx[kMaxNs];
  if (threadIdx.x < (unsigned)ns) smn[threadIdx.x] = INT32_MAX, smx[threadIdx.x] = INT32_MIN;
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    for (int s = 0; s < ns; ++s) {
      const int32_t v = states[i * ns + s];
      atomicMin(&smn[s], v);
      atomicMax(&smx[s], v);
    }
  __syncthreads();
  if (threadIdx.x < (unsigned)ns) {
    atomicMin(&mn[threadIdx.x], smn[threadIdx.x]);
    atomicMax(&mx[threadIdx.x], smx[threadIdx.x]);
  }
}

struct KeySpec {
  int nk;
  int32_t keep[kMaxNs];
  int32_t lo[kMaxNs];
  int32_t shift[kMaxNs];
};

__global__ void k_pack(int64_t n, int ns, const int32_t* states, KeySpec ks, uint64_t* keys, uint32_t* idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t k = 0;
  for (int q = 0; q < ks.nk; ++q)
    k |= (uint64_t)(uint32_t)(states[i * ns + ks.keep[q]] - ks.lo[q]) << ks.shift[q];
  keys[i] = k;
  idx[i] = (uint32_t)i;
}

__global__ void k_heads(int64_t n, const uint64_t* keys, uint32_t* head) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  head[j] = (j == 0 || keys[j] != keys[j - 1]) ? 1u : 0u;
}

// seg[g] = first sorted position of group g (seg[G] = n); first[g] = its smallest state
// index (the stable sort keeps state order inside a group); gid[g] = g.
__global__ void k_groups(int64_t n, const uint32_t* head, const uint32_t* excl, const uint32_t* idx_sorted,
                         uint32_t* seg, uint32_t* first, uint32_t* gid, uint32_t G) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j == 0) seg[G] = (uint32_t)n;
  if (j >= n || !head[j]) return;
  const uint32_t g = excl[j];
  seg[g] = (uint32_t)j;
  first[g] = idx_sorted[j];
  gid[g] = g;
}

__global__ void k_fold(uint32_t G, const uint32_t* order, const uint32_t* first_sorted, const uint32_t* seg,
                       const uint32_t* idx_sorted, const double* values, int ns, const int32_t* states, KeySpec ks,
                       int32_t* out_states, double* out_values) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= G) return;
  const uint32_t g = order[p];
  const uint32_t a = seg[g], b = seg[g + 1];
  double acc = values[idx_sorted[a]];
  for (uint32_t j = a + 1; j < b; ++j) acc += values[idx_sorted[j]];
  out_values[p] = acc;
  const int64_t f = first_sorted[p];
  for (int q = 0; q < ks.nk; ++q) out_states[p * ks.nk + q] = states[f * ns + ks.keep[q]];
}

__global__ void k_dense_last(int64_t n, int ns, const int32_t* states, const int64_t* dims, int32_t* last) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t pos = 0;
  for (int s = ns - 1; s >= 0; --s) pos = pos * dims[s] + states[i * ns + s];
  atomicMax(&last[pos], (int32_t)i);
}

__global__ void k_dense_write(int64_t n, int ns, const int32_t* states, const int64_t* dims, const int32_t* last,
                              const double* values, double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t pos = 0;
  for (int s = ns - 1; s >= 0; --s) pos = pos * dims[s] + states[i * ns + s];
  if (last[pos] == (int32_t)i) out[pos] = values[i];
}

__global__ void k_fill_i32(int64_t n, int32_t v, int32_t* out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = v;
}

__global__ void k_fill_f64(int64_t n, double* out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = 0.0;
}

// Total in a fixed order: per-block tree sums, then one block folds the partials.
__global__ void k_sum_partial(int64_t n, const double* v, double* part) {
  __shared__ double sh[kB];
  double acc = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kB;
  for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < n; i += stride) acc += v[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < (unsigned)w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_sum_final(int nb, const double* part, double* out) {
  __shared__ double sh[kB];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += kB) acc += part[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < (unsigned)w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

void minmax(const nc_fspvec_s* v, hipStream_t s, std::vector<int32_t>& mn, std::vector<int32_t>& mx) {
  const int ns = v->ns;
  mn.assign(ns, INT32_MAX);
  mx.assign(ns, INT32_MIN);
  if (v->n == 0) return;
  DevBuf<int32_t> d(2 * ns);
  NC_CHECK_HIP(hipMemcpyAsync(d.p, mn.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice, s));
  NC_CHECK_HIP(hipMemcpyAsync(d.p + ns, mx.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice, s));
  const unsigned g = std::min<unsigned>(1024, grid_for(v->n, kB));
  k_minmax<<<g, kB, 0, s>>>(v->n, ns, v->states->p, d.p, d.p + ns);
  NC_CHECK_HIP(hipGetLastError());
  int32_t* h = static_cast<int32_t*>(host_stage(0));
  NC_CHECK_HIP(hipMemcpyAsync(h, d.p, sizeof(int32_t) * 2 * ns, hipMemcpyDeviceToHost, s));
  NC_CHECK_HIP(hipStreamSynchronize(s));
  for (int k = 0; k < ns; ++k) mn[k] = h[k], mx[k] = h[ns + k];
}

int bit_width(uint32_t x) { return x ? 32 - __builtin_clz(x) : 0; }

nc_fspvec_s* marginal(const nc_fspvec_s* v, const std::vector<int32_t>& keep) {
  Stream st;
  hipStream_t s = st.s;
  const int64_t n = v->n;
  const int nk = (int)keep.size();
  auto out = std::make_unique<nc_fspvec_s>();
  out->ns = nk;
  out->states = std::make_shared<DevBuf<int32_t>>();
  NC_REQUIRE(n <= (int64_t)UINT32_MAX - 1, NC_ERR_OVERFLOW, "vector too long for the marginal");
  if (n == 0) return out.release();
  std::vector<int32_t> mn, mx;
  minmax(v, s, mn, mx);
  KeySpec ks{};
  ks.nk = nk;
  int bits = 0;
  for (int q = 0; q < nk; ++q) {
    const int sp = keep[q];
    ks.keep[q] = sp;
    ks.lo[q] = mn[sp];
    ks.shift[q] = bits;
    bits += bit_width((uint32_t)((int64_t)mx[sp] - mn[sp]));
  }
  NC_REQUIRE(bits <= 64, NC_ERR_UNSUPPORTED, "the kept coordinates need more than 64 bits of key");
  const int end_bit = std::max(bits, 1);
  DevBuf<uint64_t> keys(n), keys_sorted(n);
  DevBuf<uint32_t> idx(n), idx_sorted(n), head(n), excl(n);
  const unsigned g = grid_for(n, kB);
  k_pack<<<g, kB, 0, s>>>(n, v->ns, v->states->p, ks, keys.p, idx.p);
  NC_CHECK_HIP(hipGetLastError());
  DevBuf<unsigned char> tmp;
  size_t bytes = 0;
  NC_CHECK_HIP(rocprim::radix_sort_pairs(nullptr, bytes, keys.p, keys_sorted.p, idx.p, idx_sorted.p, (size_t)n, 0,
                                         end_bit, s));
  tmp.reserve(bytes + 16);
  NC_CHECK_HIP(rocprim::radix_sort_pairs(tmp.p, bytes, keys.p, keys_sorted.p, idx.p, idx_sorted.p, (size_t)n, 0,
                                         end_bit, s));
  k_heads<<<g, kB, 0, s>>>(n, keys_sorted.p, head.p);
  NC_CHECK_HIP(hipGetLastError());
  uint32_t G = 0;
  exclusive_scan_u32(head.p, excl.p, n, tmp, s, &G);
  DevBuf<uint32_t> seg(G + 1), first(G), gid(G), first_sorted(G), order(G);
  k_groups<<<g, kB, 0, s>>>(n, head.p, excl.p, idx_sorted.p, seg.p, first.p, gid.p, G);
  NC_CHECK_HIP(hipGetLastError());
  bytes = 0;
  NC_CHECK_HIP(rocprim::radix_sort_pairs(nullptr, bytes, first.p, first_sorted.p, gid.p, order.p, (size_t)G, 0,
                                         std::max(1, bit_width((uint32_t)(n - 1))), s));
  tmp.reserve(bytes + 16);
  NC_CHECK_HIP(rocprim::radix_sort_pairs(tmp.p, bytes, first.p, first_sorted.p, gid.p, order.p, (size_t)G, 0,
                                         std::max(1, bit_width((uint32_t)(n - 1))), s));
  out->n = G;
  if (nk) out->states->alloc((size_t)G * nk);
  out->values.alloc(G);
  k_fold<<<grid_for(G, kB), kB, 0, s>>>(G, order.p, first_sorted.p, seg.p, idx_sorted.p, v->values.p, v->ns,
                                        v->states->p, ks, out->states->p, out->values.p);
  NC_CHECK_HIP(hipGetLastError());
  NC_CHECK_HIP(hipStreamSynchronize(s));
  return out.release();
}

}  // namespace

std::shared_ptr<DevBuf<int32_t>> share_states(const int32_t* d_states, int64_t n, int ns, hipStream_t s) {
  auto st = std::make_shared<DevBuf<int32_t>>();
  if (n > 0 && ns > 0) {
    st->alloc((size_t)n * ns);
    NC_CHECK_HIP(hipMemcpyAsync(st->p, d_states, sizeof(int32_t) * n * ns, hipMemcpyDeviceToDevice, s));
  }
  return st;
}

nc_fspvec_s* fspvec_from_device(int ns, int64_t n, std::shared_ptr<DevBuf<int32_t>> states, const double* values,
                                hipStream_t src) {
  auto v = std::make_unique<nc_fspvec_s>();
  v->ns = ns;
  v->n = n;
  v->states = std::move(states);
  if (n > 0) {
    v->values.alloc(n);
    NC_CHECK_HIP(hipMemcpyAsync(v->values.p, values, sizeof(double) * n, hipMemcpyDeviceToDevice, src));
  }
  NC_CHECK_HIP(hipStreamSynchronize(src));
  return v.release();
}

}  // namespace nc

extern "C" {

int nc_fspvec_create(int32_t ns, int64_t n, const int32_t* states, const double* values, nc_fspvec* out) {
  return guard([&] {
    NC_REQUIRE(out && ns >= 0 && ns <= kMaxNs && n >= 0 && (n == 0 || ((states || ns == 0) && values)), NC_ERR_ARG,
               "bad arguments");
    auto v = std::make_unique<nc_fspvec_s>();
    v->ns = ns;
    v->n = n;
    v->states = std::make_shared<DevBuf<int32_t>>();
    Stream st;
    hipStream_t s = st.s;
    if (n > 0) {
      if (ns > 0) {
        v->states->alloc((size_t)n * ns);
        NC_CHECK_HIP(hipMemcpyAsync(v->states->p, states, sizeof(int32_t) * n * ns, hipMemcpyHostToDevice, s));
      }
      v->values.alloc(n);
      NC_CHECK_HIP(hipMemcpyAsync(v->values.p, values, sizeof(double) * n, hipMemcpyHostToDevice, s));
    }
    NC_CHECK_HIP(hipStreamSynchronize(s));
    *out = v.release();
  });
}

int nc_fspvec_destroy(nc_fspvec v) {
  return guard([&] { delete v; });
}

int nc_fspvec_size(nc_fspvec v, int64_t* n, int32_t* ns) {
  return guard([&] {
    NC_REQUIRE(v, NC_ERR_ARG, "null vector");
    if (n) *n = v->n;
    if (ns) *ns = v->ns;
  });
}

int nc_fspvec_get(nc_fspvec v, int32_t* states, double* values) {
  return guard([&] {
    NC_REQUIRE(v, NC_ERR_ARG, "null vector");
    Stream st;
    hipStream_t s = st.s;
    if (v->n > 0) {
      if (states && v->ns > 0) copy_d2h(states, v->states->p, sizeof(int32_t) * v->n * v->ns, s);
      if (values) copy_d2h(values, v->values.p, sizeof(double) * v->n, s);
    }
    NC_CHECK_HIP(hipStreamSynchronize(s));
  });
}

int nc_fspvec_sum(nc_fspvec v, double* total) {
  return guard([&] {
    NC_REQUIRE(v && total, NC_ERR_ARG, "null argument");
    *total = 0.0;
    if (v->n == 0) return;
    Stream st;
    hipStream_t s = st.s;
    const int nb = (int)std::min<int64_t>(1024, grid_for(v->n, kB));
    DevBuf<double> part(nb + 1);
    k_sum_partial<<<nb, kB, 0, s>>>(v->n, v->values.p, part.p);
    k_sum_final<<<1, kB, 0, s>>>(nb, part.p, part.p + nb);
    NC_CHECK_HIP(hipGetLastError());
    double* h = static_cast<double*>(host_stage(0));
    NC_CHECK_HIP(hipMemcpyAsync(h, part.p + nb, sizeof(double), hipMemcpyDeviceToHost, s));
    NC_CHECK_HIP(hipStreamSynchronize(s));
    *total = h[0];
  });
}

int nc_fspvec_marginal(nc_fspvec v, int32_t n_dims, const int32_t* dims, nc_fspvec* out) {
  return guard([&] {
    NC_REQUIRE(v && out && n_dims >= 1 && dims, NC_ERR_ARG, "bad arguments");
    std::vector<bool> drop(v->ns, false);
    for (int k = 0; k < n_dims; ++k) {
      NC_REQUIRE(dims[k] >= 0 && dims[k] < v->ns, NC_ERR_ARG,
                 "Input dimensions must be between 1 and " + std::to_string(v->ns) + ".");
      drop[dims[k]] = true;
    }
    std::vector<int32_t> keep;  // setdiff(1:NS, dims), ascending
    for (int sp = 0; sp < v->ns; ++sp)
      if (!drop[sp]) keep.push_back(sp);
    *out = marginal(v, keep);
  });
}

int nc_fspvec_dense(nc_fspvec v, int64_t* bounds, int64_t capacity, double* out) {
  return guard([&] {
    NC_REQUIRE(v && bounds, NC_ERR_ARG, "null argument");
    NC_REQUIRE(v->n > 0, NC_ERR_ARG, "Cannot construct dense array from empty FspVector instance.");
    Stream st;
    hipStream_t s = st.s;
    std::vector<int32_t> mn, mx;
    minmax(v, s, mn, mx);
    int64_t total = 1;
    for (int sp = 0; sp < v->ns; ++sp) {
      NC_REQUIRE(mn[sp] >= 0, NC_ERR_ARG, "dense conversion needs non-negative states");
      bounds[sp] = (int64_t)mx[sp] + 1;
      NC_REQUIRE(total <= ((int64_t)1 << 40) / bounds[sp], NC_ERR_OVERFLOW, "dense array too large");
      total *= bounds[sp];
    }
    if (!out) return;
    NC_REQUIRE(capacity >= total, NC_ERR_ARG, "output buffer smaller than the dense array");
    NC_REQUIRE(v->n <= INT32_MAX, NC_ERR_OVERFLOW, "vector too long for dense conversion");
    DevBuf<double> dense(total);
    DevBuf<int32_t> last(total);
    DevBuf<int64_t> ddims(std::max(v->ns, 1));
    NC_CHECK_HIP(hipMemcpyAsync(ddims.p, bounds, sizeof(int64_t) * v->ns, hipMemcpyHostToDevice, s));
    const unsigned gf = std::min<unsigned>(4096, grid_for(total, kB));
    k_fill_i32<<<gf, kB, 0, s>>>(total, -1, last.p);
    k_fill_f64<<<gf, kB, 0, s>>>(total, dense.p);
    const unsigned g = grid_for(v->n, kB);
    k_dense_last<<<g, kB, 0, s>>>(v->n, v->ns, v->states->p, ddims.p, last.p);
    k_dense_write<<<g, kB, 0, s>>>(v->n, v->ns, v->states->p, ddims.p, last.p, v->values.p, dense.p);
    NC_CHECK_HIP(hipGetLastError());
    NC_CHECK_HIP(hipMemcpyAsync(out, dense.p, sizeof(double) * total, hipMemcpyDeviceToHost, s));
    NC_CHECK_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
