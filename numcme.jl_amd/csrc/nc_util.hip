// Library-wide utilities: error state, version/device queries, the rocPRIM prefix sum
// used by the state space, and the host dense matrix exponential used by the Krylov
// integrator.
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "nc_internal.h"

namespace nc {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, DevBuf<unsigned char>& tmp, hipStream_t s,
                        uint32_t* total_host) {
  if (n <= 0) {
    if (total_host) *total_host = 0;
    return;
  }
  size_t bytes = 0;
  NC_CHECK_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(), s));
  tmp.reserve(bytes + 16);
  NC_CHECK_HIP(rocprim::exclusive_scan(tmp.p, bytes, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(), s));
  if (total_host) {
    uint32_t* h = static_cast<uint32_t*>(host_stage(0));
    NC_CHECK_HIP(hipMemcpyAsync(h, in + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    NC_CHECK_HIP(hipMemcpyAsync(h + 1, out + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    NC_CHECK_HIP(hipStreamSynchronize(s));
    *total_host = h[0] + h[1];
  }
}

// ---- dense matrix exponential (host, FP64) --------------------------------------------
// Scaling and squaring with the [13/13] Pade approximant and the theta_13 threshold of
// N. J. Higham, "The scaling and squaring method for the matrix exponential revisited",
// SIAM J. Matrix Anal. Appl. 26(4), 2005. The Krylov integrator calls it at least once per
// step on its (m+2)x(m+2) Hessenberg matrix, between a device synchronisation and the next
// launch, so it is on the solve's critical path: matrices of at most kExpmMax rows are
// padded to a multiple of 8 and multiplied with a row of the product held in vector
// registers (AVX2 and FMA when the host has them, checked once; plain vectors otherwise).
// The Krylov step needs only the first column, which saves the last squaring.
namespace {

constexpr int kExpmMax = 64;
namespace expm_avx2 {
#pragma clang attribute push(__attribute__((target("avx2,fma"))), apply_to = function)
#include "nc_expm_impl.h"
#pragma clang attribute pop
}  // namespace expm_avx2

namespace expm_base {
#include "nc_expm_impl.h"
}  // namespace expm_base

bool host_has_avx2() {
  static const bool yes = __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
  return yes;
}

// the plain algorithm for matrices above kExpmMax rows (not used by the integrators)
void matmul_plain(int n, const double* a, const double* b, double* c) {
  std::fill(c, c + (size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) {
      const double aik = a[i * n + k];
      if (aik == 0.0) continue;
      for (int j = 0; j < n; ++j) c[i * n + j] += aik * b[k * n + j];
    }
}

void expm_large(int n, const double* a, double* out) {
  static const double b[14] = {64764752532480000.0, 32382376266240000.0, 7771770303897600.0, 1187353796428800.0,
                               129060195264000.0,   10559470521600.0,    670442572800.0,     33522128640.0,
                               1323241920.0,        40840800.0,          960960.0,           16380.0,
                               182.0,               1.0};
  const double theta13 = 5.371920351148152;
  const size_t nn = (size_t)n * n;
  double norm1 = 0.0;
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += std::fabs(a[i * n + j]);
    norm1 = std::max(norm1, s);
  }
  int sq = 0;
  if (norm1 > theta13) sq = std::max(0, (int)std::ceil(std::log2(norm1 / theta13)));
  const double scale = std::ldexp(1.0, -sq);
  std::vector<double> A(nn), A2(nn), A4(nn), A6(nn), U(nn), V(nn), T(nn), W(nn);
  for (size_t k = 0; k < nn; ++k) A[k] = a[k] * scale;
  matmul_plain(n, A.data(), A.data(), A2.data());
  matmul_plain(n, A2.data(), A2.data(), A4.data());
  matmul_plain(n, A4.data(), A2.data(), A6.data());
  for (size_t k = 0; k < nn; ++k) T[k] = b[13] * A6[k] + b[11] * A4[k] + b[9] * A2[k];
  matmul_plain(n, A6.data(), T.data(), W.data());
  for (size_t k = 0; k < nn; ++k) W[k] += b[7] * A6[k] + b[5] * A4[k] + b[3] * A2[k];
  for (int i = 0; i < n; ++i) W[(size_t)i * n + i] += b[1];
  matmul_plain(n, A.data(), W.data(), U.data());
  for (size_t k = 0; k < nn; ++k) T[k] = b[12] * A6[k] + b[10] * A4[k] + b[8] * A2[k];
  matmul_plain(n, A6.data(), T.data(), V.data());
  for (size_t k = 0; k < nn; ++k) V[k] += b[6] * A6[k] + b[4] * A4[k] + b[2] * A2[k];
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] += b[0];
  std::vector<double> Q(nn), P(nn);
  for (size_t k = 0; k < nn; ++k) Q[k] = V[k] - U[k], P[k] = V[k] + U[k];
  for (int col = 0; col < n; ++col) {
    int piv = col;
    double best = std::fabs(Q[col * n + col]);
    for (int r = col + 1; r < n; ++r)
      if (std::fabs(Q[r * n + col]) > best) best = std::fabs(Q[r * n + col]), piv = r;
    if (best == 0.0) throw Error(NC_ERR_NUMERIC, "singular Pade denominator in expm");
    if (piv != col)
      for (int j = 0; j < n; ++j) std::swap(Q[col * n + j], Q[piv * n + j]), std::swap(P[col * n + j], P[piv * n + j]);
    const double inv = 1.0 / Q[col * n + col];
    for (int r = col + 1; r < n; ++r) {
      const double f = Q[r * n + col] * inv;
      if (f == 0.0) continue;
      for (int j = col; j < n; ++j) Q[r * n + j] -= f * Q[col * n + j];
      for (int j = 0; j < n; ++j) P[r * n + j] -= f * P[col * n + j];
    }
  }
  for (int col = n - 1; col >= 0; --col) {
    const double inv = 1.0 / Q[col * n + col];
    for (int j = 0; j < n; ++j) P[col * n + j] *= inv;
    for (int r = 0; r < col; ++r) {
      const double f = Q[r * n + col];
      if (f == 0.0) continue;
      for (int j = 0; j < n; ++j) P[r * n + j] -= f * P[col * n + j];
    }
  }
  for (int k = 0; k < sq; ++k) {
    matmul_plain(n, P.data(), P.data(), T.data());
    std::swap(P, T);
  }
  std::copy(P.begin(), P.end(), out);
}

// one work area per host thread, on the heap (the arrays are 256 KB in all); both kernel
// sets use the same layout
struct alignas(64) ExpmArea {
  unsigned char bytes[sizeof(expm_base::ExpmWork)];
};
static_assert(sizeof(expm_base::ExpmWork) == sizeof(expm_avx2::ExpmWork), "work area layouts differ");

void expm_small(int n, const double* a, double* out, bool first_column) {
  static thread_local std::unique_ptr<ExpmArea> area;
  if (!area) area = std::make_unique<ExpmArea>();
  if (host_has_avx2())
    expm_avx2::expm_small(n, a, out, first_column, *reinterpret_cast<expm_avx2::ExpmWork*>(area.get()));
  else
    expm_base::expm_small(n, a, out, first_column, *reinterpret_cast<expm_base::ExpmWork*>(area.get()));
}

}  // namespace

void expm_dense(int n, const double* a, double* out) {
  if (n > kExpmMax)
    expm_large(n, a, out);
  else
    expm_small(n, a, out, false);
}

void expm_first_column(int n, const double* a, double* out) {
  if (n > kExpmMax) {
    std::vector<double> full((size_t)n * n);
    expm_large(n, a, full.data());
    for (int i = 0; i < n; ++i) out[i] = full[(size_t)i * n];
    return;
  }
  expm_small(n, a, out, true);
}

}  // namespace nc

using namespace nc;

extern "C" {

int nc_last_error(char* buf, size_t len) {
  if (!buf || len == 0) return NC_ERR_ARG;
  size_t k = std::min(len - 1, g_last_error.size());
  std::memcpy(buf, g_last_error.data(), k);
  buf[k] = '\0';
  return NC_OK;
}

int nc_memory_trim(void) {
  return guard([&] {
    NC_CHECK_HIP(hipDeviceSynchronize());
    pool_trim();
  });
}

int nc_version(int32_t* major, int32_t* minor) {
  if (major) *major = 0;
  if (minor) *minor = 1;
  return NC_OK;
}

int nc_device_name(char* buf, size_t len) {
  return guard([&] {
    NC_REQUIRE(buf && len, NC_ERR_ARG, "null buffer");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw Error(NC_ERR_NO_DEVICE, "no HIP device available");
    int dev = 0;
    NC_CHECK_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    NC_CHECK_HIP(hipGetDeviceProperties(&prop, dev));
    std::string s = std::string(prop.name) + " (" + prop.gcnArchName + ", " + std::to_string(prop.multiProcessorCount) + " CUs)";
    size_t k = std::min(len - 1, s.size());
    std::memcpy(buf, s.data(), k);
    buf[k] = '\0';
  });
}

int nc_expm_dense(int32_t n, const double* a, double* out) {
  return guard([&] {
    NC_REQUIRE(n > 0 && a && out, NC_ERR_ARG, "bad expm arguments");
    expm_dense(n, a, out);
  });
}

}  // extern "C"
